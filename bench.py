"""Benchmark: plain GPU mismatch scan over a 3.1 Gbp synthetic genome (one MI355X).

Workload (the configuration the headline metric is quoted on): 24 chromosomes
of 50-250 Mbp rescaled to 3.1 Gbp, 41% GC, 5% N in gaps of 10 kbp-1 Mbp;
100 random guides for each length {20, 25} and PAM {NRG, NNGRRT (3'),
TTTN (5')}, each run at M in {3, 6} on both strands, 1,000 planted near-hits
per guide.  That is 12 reference-format runs (one filter, length and M per
run), each a separate libcof context with the genome resident in HBM.

A step is the 12 searches (kernels, D2H of hits, conversion to final hits).
value = guides x genome bp (N included) x 12 runs / step time, in
guide-Gbp/s.  The roofline fields price the plain-scan kernel by the
reference's brute-force op count per guide-bp, 2*(3*ceil(L/8)+1) int32 ops,
against gfx950's int32 VALU peak; the kernel time comes from HIP events on the
library's stream.  cpu_baseline times the CPU oracle (a restatement of the
reference's CPU path) on a bounded sample of the same workload.
"""
import argparse
import csv
import glob
import hashlib
import json
import math
import os
import random
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cas-offinder-rust-old_amd"))

import cofpy  # noqa: E402

INT32_PEAK = 256 * 64 * 2.4e9       # gfx950: 256 CUs x 64 lanes x 2.4 GHz int32 VALU ops/s
HBM_PEAK = 8.0e12
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4   # wave64 VALU instructions/s: 4 SIMDs per CU, 4 cycles each
MFMA_FP4_PEAK = 10.1e15                # dense fp4 matrix-core ops/s
MFMA_FP4_OPS = 2 * 16 * 16 * 128      # ops of one v_mfma_f32_16x16x128_f8f6f4
DIGESTS = os.path.join(ROOT, "tests", "golden", "headline_3.1Gbp_digest.json")
COUNTERS = os.path.join(ROOT, "profiles", "plain_scan_counters.json")
PLAIN_SRC = os.path.join(ROOT, "cas-offinder-rust-old_amd", "csrc", "plain.hip")
PAMS = [("NRG", True), ("NNGRRT", True), ("TTTN", False)]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def file_sha256(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def rows_digest(path):
    """(row count, SHA-256 of the sorted data rows, each followed by a newline) of a reference TSV."""
    rows = cofpy.read_rows(path)
    h = hashlib.sha256()
    for r in rows:
        h.update(r.encode())
        h.update(b"\n")
    return len(rows), h.hexdigest()


def measured_counters():
    """Counter figures of plain_scan_kernel from the committed profile summary, or None when the kernel
    source changed since they were measured (they would describe another kernel)."""
    if not os.path.exists(COUNTERS):
        return None, "no committed counter summary"
    with open(COUNTERS) as f:
        doc = json.load(f)
    if doc.get("plain_hip_sha256") != file_sha256(PLAIN_SRC):
        return None, "profiles/plain_scan_counters.json was measured on another plain.hip"
    return doc, "profiles/plain_scan_counters.json (measured on this plain.hip)"


# ----------------------------------------------------------------------------- rocprofv3 summaries
def _csv_rows(d, suffix):
    out = []
    for path in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(path, newline="") as f:
            out += list(csv.DictReader(f))
    return out


def _dispatches(d):
    """dispatch id -> {"name", "ns", counter: value} from a rocprofv3 --pmc csv folder."""
    disp = {}
    for r in _csv_rows(d, "counter_collection.csv"):
        k = int(r["Dispatch_Id"])
        e = disp.setdefault(k, {"name": r["Kernel_Name"], "ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"]),
                                "vgpr": r.get("Arch_VGPR_Count") or r.get("VGPR_Count"),
                                "lds": r.get("LDS_Block_Size") or r.get("Group_Segment_Size")})
        e[r["Counter_Name"]] = e.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return disp


def _last_plain(disp, n):
    ks = sorted(k for k, e in disp.items() if "plain_scan_kernel" in e["name"])
    if len(ks) < n:
        raise SystemExit(f"expected at least {n} plain_scan_kernel dispatches, found {len(ks)}")
    return [disp[k] for k in ks[-n:]]


def _copy_factor(disp, counter, nbytes):
    """bytes per counter unit: the known size of the calibration copy over what the counter reads for it."""
    ks = [k for k, e in disp.items() if counter in e and ("copy" in e["name"].lower() or "elementwise" in e["name"].lower())]
    if not ks:
        raise SystemExit(f"no calibration copy kernel with {counter} found")
    return nbytes / disp[max(ks, key=lambda k: disp[k][counter])][counter]


def summarize_counters(args):
    """profiles/plain_scan_counters.json + a markdown table from rocprofv3 runs of this script with
    --no-overlap --warmup 1 --steps 1 --calibrate-copy --manifest M (one pass per folder)."""
    with open(args.manifest) as f:
        man = json.load(f)
    runs = man["runs"]            # [[name, launches, hits, L, genome_nucl]] in dispatch order
    n = sum(r[1] for r in runs)
    names = [r[0] for r in runs for _ in range(r[1])]
    per = [{"run": nm} for nm in names]
    out = {"plain_hip_sha256": file_sha256(PLAIN_SRC), "workload": man["workload"], "units": {}}
    fetch = _dispatches(args.counters_from[0])
    fdisp = _last_plain(fetch, n)
    ffac = _copy_factor(fetch, "FETCH_SIZE", man["copy_bytes"])
    write = _dispatches(args.counters_from[1])
    wdisp = _last_plain(write, n)
    wfac = _copy_factor(write, "WRITE_SIZE", man["copy_bytes"])
    sq = _last_plain(_dispatches(args.counters_from[2]), n)
    mf = _last_plain(_dispatches(args.counters_from[3]), n)
    for i, p in enumerate(per):
        t = sq[i]["ns"] * 1e-9
        p["ms_pmc"] = round(sq[i]["ns"] / 1e6, 3)
        p["fetch_gb"] = round(fdisp[i]["FETCH_SIZE"] * ffac / 1e9, 4)
        p["write_gb"] = round(wdisp[i]["WRITE_SIZE"] * wfac / 1e9, 4)
        p["valu_issue"] = round(sq[i]["SQ_INSTS_VALU"] / (t * VALU_ISSUE_PEAK), 3)
        p["valu_per_wave"] = round(sq[i]["SQ_INSTS_VALU"] / sq[i]["SQ_WAVES"], 1)
        p["lds_per_wave"] = round(sq[i]["SQ_INSTS_LDS"] / sq[i]["SQ_WAVES"], 1)
        p["bank_conflict_cycles_per_lds"] = round(sq[i]["SQ_LDS_BANK_CONFLICT"] / max(sq[i]["SQ_INSTS_LDS"], 1), 3)
        p["clock_ghz"] = round(sq[i]["GRBM_GUI_ACTIVE"] / 8 / t / 1e9, 2)
        mt = mf[i]["ns"] * 1e-9
        # every matrix-core instruction of the kernel is one fp4 16x16x128 MFMA
        p["mfma_fp4_tops"] = round(mf[i]["SQ_INSTS_MFMA"] * MFMA_FP4_OPS / mt / 1e12, 1)
        p["mfma_per_wave"] = round(mf[i]["SQ_INSTS_MFMA"] / max(sq[i]["SQ_WAVES"], 1), 1)
        clk = mf[i]["GRBM_GUI_ACTIVE"] / 8
        p["mfma_busy"] = round(mf[i]["SQ_VALU_MFMA_BUSY_CYCLES"] / (clk * 1024), 3)
        p["rocprof_vgpr_count"] = sq[i]["vgpr"]
        p["lds_bytes"] = sq[i]["lds"]
    tot = sum(p["ms_pmc"] for p in per)
    w = lambda k: round(sum(p[k] * p["ms_pmc"] for p in per) / tot, 3)
    out.update({"runs": per, "traffic_bytes_per_launch": round(sum((p["fetch_gb"] + p["write_gb"]) * 1e9 for p in per) / n),
                "fetch_factor": round(ffac, 4), "write_factor": round(wfac, 4),
                "valu_issue_frac": w("valu_issue"), "mfma_busy_frac": w("mfma_busy"),
                "mfma_fp4_tops": round(sum(p["mfma_fp4_tops"] * p["ms_pmc"] for p in per) / tot, 1),
                "bank_conflict_cycles_per_lds": w("bank_conflict_cycles_per_lds")})
    out["units"] = {"fetch_gb": "FETCH_SIZE x fetch_factor (bytes a 256 MiB copy reads per unit)",
                    "write_gb": "WRITE_SIZE x write_factor", "valu_issue": "SQ_INSTS_VALU / (time x 614.4 G/s)",
                    "mfma_busy": "SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / 8 XCDs x 1024 SIMDs)",
                    "mfma_fp4_tops": "SQ_INSTS_MFMA x 65,536 (ops of one fp4 16x16x128 MFMA) / time",
                    "weights": "time-weighted means over the runs"}
    with open(args.counters_out or COUNTERS, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    cols = ["run", "ms_pmc", "fetch_gb", "write_gb", "valu_per_wave", "valu_issue", "lds_per_wave",
            "bank_conflict_cycles_per_lds", "mfma_per_wave", "mfma_fp4_tops", "mfma_busy", "clock_ghz", "rocprof_vgpr_count",
            "lds_bytes"]
    print("| " + " | ".join(cols) + " |")
    print("|" + "---|" * len(cols))
    for p in per:
        print("| " + " | ".join(str(p[c]) for c in cols) + " |")
    print(json.dumps({k: v for k, v in out.items() if k != "runs"}))


def summarize_bulge_counters(args):
    """profiles/bulge_counters.json: the bulge workload's filter and traceback kernels summed over their
    dispatches (rocprofv3 --pmc runs of --workload bulge --schedule sync --calibrate-copy), each placed on the
    HBM and VALU-issue rooflines."""
    fetch, write, sq, mf = (_dispatches(d) for d in args.counters_from)
    ffac = _copy_factor(fetch, "FETCH_SIZE", 256 << 20)
    wfac = _copy_factor(write, "WRITE_SIZE", 256 << 20)
    out = {"bulge_hip_sha256": file_sha256(os.path.join(ROOT, "cas-offinder-rust-old_amd", "csrc", "bulge.hip")),
           "workload": "bench.py --workload bulge (config 3: 3.1 Gbp repeat-rich, 100 guides NRG M4 D2 R2), "
                       "synchronous rounds", "kernels": {}}
    for key in ("bulge_filter_kernel", "bulge_screen_kernel", "bulge_verify_kernel"):
        def tot(disp, c):
            return sum(e.get(c, 0.0) for e in disp.values() if key in e["name"])
        ds = [e for e in sq.values() if key in e["name"]]
        if not ds:
            continue
        t = sum(e["ns"] for e in ds) * 1e-9
        tm = sum(e["ns"] for e in mf.values() if key in e["name"]) * 1e-9
        fb, wb = tot(fetch, "FETCH_SIZE") * ffac, tot(write, "WRITE_SIZE") * wfac
        tf = sum(e["ns"] for e in fetch.values() if key in e["name"]) * 1e-9
        k = {"dispatches": len(ds), "ms_pmc": round(t * 1e3, 2),
             "fetch_gb": round(fb / 1e9, 3), "write_gb": round(wb / 1e9, 3),
             "hbm_gb_s": round((fb + wb) / max(tf, 1e-12) / 1e9, 1),
             "hbm_frac": round((fb + wb) / max(tf, 1e-12) / HBM_PEAK, 4),
             "valu_issue": round(tot(sq, "SQ_INSTS_VALU") / (t * VALU_ISSUE_PEAK), 3),
             "salu_per_valu": round(tot(sq, "SQ_INSTS_SALU") / max(tot(sq, "SQ_INSTS_VALU"), 1), 3),
             "valu_per_wave": round(tot(sq, "SQ_INSTS_VALU") / max(tot(sq, "SQ_WAVES"), 1), 1),
             "lds_per_valu": round(tot(sq, "SQ_INSTS_LDS") / max(tot(sq, "SQ_INSTS_VALU"), 1), 3),
             "bank_conflict_cycles_per_lds": round(tot(sq, "SQ_LDS_BANK_CONFLICT") / max(tot(sq, "SQ_INSTS_LDS"), 1), 3),
             "clock_ghz": round(tot(sq, "GRBM_GUI_ACTIVE") / 8 / t / 1e9, 2),
             "rocprof_vgpr_count": ds[0]["vgpr"], "lds_bytes": ds[0]["lds"]}
        if tm:
            k["valu_busy"] = round(tot(mf, "SQ_ACTIVE_INST_VALU") / (tot(mf, "GRBM_GUI_ACTIVE") / 8 * 1024), 3) \
                if tot(mf, "SQ_ACTIVE_INST_VALU") else None
        out["kernels"][key] = k
    out["units"] = {"fetch_gb": "FETCH_SIZE x fetch_factor over all dispatches", "hbm_frac": "(fetch + write) / "
                    "FETCH run's dispatch time / 8 TB/s", "valu_issue": "SQ_INSTS_VALU / (time x 614.4 G/s)"}
    out["fetch_factor"], out["write_factor"] = round(ffac, 4), round(wfac, 4)
    with open(args.counters_out or os.path.join(ROOT, "profiles", "bulge_counters.json"), "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out, indent=1))


def chrom_lengths(total, n, seed):
    rng = random.Random(seed)
    raw = [rng.uniform(50e6, 250e6) for _ in range(n)]
    scale = total / sum(raw)
    lens = [int(x * scale) for x in raw]
    lens[-1] += total - sum(lens)
    return lens


def make_queries(seed):
    rng = random.Random(seed)
    runs = []
    for glen in (20, 25):
        for pam, three in PAMS:
            guides = ["".join(rng.choice("ACGT") for _ in range(glen)) for _ in range(100)]
            if three:
                filt, pats = "N" * glen + pam, [g + "N" * len(pam) for g in guides]
                plants = [(g + pam, 0, glen, 6, 1000) for g in guides]
            else:
                filt, pats = pam + "N" * glen, ["N" * len(pam) + g for g in guides]
                plants = [(pam + g, len(pam), glen, 6, 1000) for g in guides]
            for mm in (3, 6):
                runs.append({"name": f"{pam}-{glen}-M{mm}", "filter": filt, "pats": pats, "mm": mm,
                             "L": len(filt), "plants": plants if mm == 3 else []})
    return runs


def ops_per_guide_bp(L):
    return 2 * (3 * math.ceil(L / 8) + 1)


def cpu_baseline(run, sample_bp, seed, threads):
    """CPU oracle on a bounded sample: same genome model, one run's 100 guides."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    sg = cofpy.SynthGenome(seed + 17, [sample_bp], n_frac=0.05, gap=(10_000, 1_000_000), plants=run["plants"][:10])
    with tempfile.TemporaryDirectory() as d:
        fa = os.path.join(d, "s.fa")
        sg.write_fasta(fa)
        in_text = f"{fa}\n{run['filter']}\n" + "".join(f"{p} {run['mm']}\n" for p in run["pats"])
        n, st = oracle.run(in_text, os.path.join(d, "o.tsv"), batch_nucl=1 << 22, threads=threads)
    secs = st["search_seconds"]
    return {"value": round(len(run["pats"]) * sample_bp / secs / 1e9, 3), "unit": "guide-Gbp/s",
            "cores": threads, "kind": "port",
            "sample": f"{run['name']}: 100 guides x {sample_bp / 1e6:.0f} Mbp synthetic, 4 Mnt batches, "
                      f"search-only time {secs:.2f} s, {n} rows"}


def calibration_copy():
    """A 256 MiB device-to-device copy kernel: the known size the traffic counters are scaled by."""
    import torch
    a = torch.empty((256 << 20) // 4, dtype=torch.int32, device="cuda").random_()
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    log("calibration copy of 256 MiB done")


def bulge_workload(args):
    """Bulge search: 3.1 Gbp layout with 10% repeat copies, 100 guides, NRG, M4 D2 R2."""
    if args.calibrate_copy:
        calibration_copy()
    total = int(args.genome_bp)
    lens = chrom_lengths(total, 24, args.seed)
    rng = random.Random(args.seed + 5)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(args.bulge_guides)]
    plants = [(g + "NRG", 0, 20, 4, 100) for g in guides]
    t = time.time()
    sg = cofpy.SynthGenome(args.seed + 3, lens, n_frac=0.05, gap=(10_000, 1_000_000),
                           repeats=(0.10, 20, 300, 0.10), plants=plants)
    log(f"repeat-rich genome {total / 1e9:.2f} Gbp generated in {time.time() - t:.1f} s")
    pats = [g + "NNN" for g in guides]
    n = len(pats)
    s = cofpy.Search(pats, ["NNNNNNNNNNNNNNNNNNNNNRG"] * n, [4] * n, [2] * n, [2] * n, count_only=True,
                     flags=schedule_flags(args))
    s.load_synth(sg)
    apply_tuning(s, args)
    for w in range(args.warmup):
        s.search()
        log(f"warmup {w + 1} done")
    times, filt, ver = [], 0.0, 0.0
    for k in range(args.steps):
        t0 = time.perf_counter()
        s.search()
        times.append(time.perf_counter() - t0)
        st = s.stats()
        filt += st.kernel_ms_filter
        ver += st.kernel_ms_verify
        log(f"step {k + 1}: {times[-1]:.3f} s, candidates {st.n_candidates} (traced {st.n_traced}), hits {st.bulge_hits}, "
            f"bulge path {st.bulge_ms:.1f} ms ({st.n_pipeline_rounds} pipeline rounds, {st.n_redo_rounds} redone; "
            f"synchronous rounds: filter {st.kernel_ms_filter:.1f} ms, verify {st.kernel_ms_verify:.1f} ms), "
            f"dfs steps {st.verify_steps} ({st.verify_steps / max(st.n_traced, 1):.1f}/traced cand), "
            f"lane use {st.verify_busy / max(st.verify_slots, 1):.3f}")
    st = s.stats()
    step = sum(times) / len(times)
    units = n * st.genome_nucl
    out = {"metric": "guide-Gbp/s at bit-exact hits (bulge search)",
           "value": round(units / step / 1e9, 3), "unit": "guide-Gbp/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(step * 1e3, 1), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32/u64 (bit-parallel Myers, u8 DP bound)", "data": "synthetic",
           "config": {"workload": f"bulge search: {total / 1e9:.2f} Gbp synthetic, 10% copies of 20 x 300-bp repeats "
                                  f"at 10% divergence, {n} guides NRG, M4 D2 R2, count+checksum sink",
                      "genome_bp": st.genome_nucl},
           "candidates": st.n_candidates, "traced": st.n_traced, "bulge_hits": st.bulge_hits,
           "checksum": st.bulge_checksum, "dfs_steps": st.verify_steps, "schedule": args.schedule, "tuning": s.tuning(), "pipeline_rounds": st.n_pipeline_rounds,
           "graph_instantiations_last_step": st.graph_instantiations, "graph_updates_last_step": st.graph_updates,
           "kernel_ms_filter": round(filt / args.steps, 1), "kernel_ms_verify": round(ver / args.steps, 1)}
    s.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = bulge_cpu_baseline(args, guides, plants)
        log(f"cpu baseline: {out['cpu_baseline']}")
    print(json.dumps(out), flush=True)


def costmodel_workload(args):
    """Cost-model work order for the bulge search: train on seeded genomes, compare wall time of orders.

    Training rows: every 64 kbp chunk of 20 repeat-rich genomes (seeds 1001-1020,
    --train-bp each) searched with the bulge query below; target log(1 + DFS
    transitions of the chunk).  Evaluation: the bulge workload's genome and query
    under static genome order, an untrained (random) model and the trained model.
    """
    import numpy as np
    import costmodel
    rng = random.Random(args.seed + 5)
    guides = ["".join(rng.choice("ACGT") for _ in range(20)) for _ in range(args.bulge_guides)]
    plants = [(g + "NRG", 0, 20, 4, 100) for g in guides]
    pats = [g + "NNN" for g in guides]
    n = len(pats)
    filt = ["NNNNNNNNNNNNNNNNNNNNNRG"] * n
    s = cofpy.Search(pats, filt, [4] * n, [2] * n, [2] * n, count_only=True, flags=schedule_flags(args))
    if args.tune:
        s.set_tuning(**{k: int(v) for k, v in (t.split("=") for t in args.tune)})
    xs, ts = [], []
    t0 = time.time()
    train_bp = int(args.train_bp)
    for k, seed in enumerate(range(1001, 1001 + args.train_genomes)):
        lens = chrom_lengths(train_bp, 24, seed)
        sg = cofpy.SynthGenome(seed, lens, n_frac=0.05, gap=(10_000, 1_000_000), repeats=(0.10, 20, 300, 0.10),
                               plants=plants)
        s.load_synth(sg)
        x, t = costmodel.collect(s)
        xs.append(x)
        ts.append(t)
        log(f"training genome {k + 1}/{args.train_genomes} (seed {seed}): {len(x)} chunks, "
            f"{time.time() - t0:.1f} s so far")
    X, T = np.concatenate(xs), np.concatenate(ts)
    t1 = time.time()
    trained, loss = costmodel.fit(X, T, steps=args.train_steps)
    pred = trained.forward(X)
    r2 = 1 - float(np.mean((pred - T) ** 2) / np.var(T))
    log(f"trained on {len(X)} rows in {time.time() - t1:.2f} s: standardised loss {loss:.4f}, R^2 {r2:.3f}")
    total = int(args.genome_bp)
    lens = chrom_lengths(total, 24, args.seed)
    sg = cofpy.SynthGenome(args.seed + 3, lens, n_frac=0.05, gap=(10_000, 1_000_000), repeats=(0.10, 20, 300, 0.10),
                           plants=plants)
    s.load_synth(sg)
    xe, te = costmodel.collect(s)
    pe = trained.forward(xe)
    r2e = 1 - float(np.mean((pe - te) ** 2) / np.var(te))
    log(f"evaluation genome {total / 1e9:.2f} Gbp: held-out R^2 {r2e:.3f}")
    untrained = cofpy.Mlp(seed=7)
    res = {}
    for name, m in (("static", None), ("untrained", untrained), ("trained", trained)):
        s.set_cost_model(m)
        times = []
        for k in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            s.search()
            dt = time.perf_counter() - t0
            st = s.stats()
            if k >= args.warmup:
                times.append(dt)
            log(f"{name} {'warmup' if k < args.warmup else 'step'}: {dt:.3f} s, hits {st.bulge_hits}, "
                f"rounds {st.n_pipeline_rounds}, redone {st.n_redo_rounds}")
        res[name] = {"seconds": round(sum(times) / len(times), 3), "bulge_hits": st.bulge_hits,
                     "checksum": st.bulge_checksum}
    out = {"metric": "bulge search wall time by work order (cost model)", "value": res["trained"]["seconds"],
           "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(res["trained"]["seconds"] * 1e3, 1), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32 (MLP), u32/u64 (search)", "data": "synthetic",
           "config": {"workload": f"bulge search {total / 1e9:.2f} Gbp repeat-rich synthetic, {n} guides NRG M4 D2 R2; "
                                  f"model trained on {args.train_genomes} x {train_bp / 1e6:.0f} Mbp genomes",
                      "genome_bp": st.genome_nucl},
           "orders": res, "speedup_vs_static": round(res["static"]["seconds"] / res["trained"]["seconds"], 3),
           "identical": len({(v["bulge_hits"], v["checksum"]) for v in res.values()}) == 1,
           "train_rows": len(X), "train_r2": round(r2, 3), "heldout_r2": round(r2e, 3)}
    s.close()
    print(json.dumps(out), flush=True)


def apply_tuning(s, args):
    if args.tune:
        s.set_tuning(**{k: int(v) for k, v in (t.split("=") for t in args.tune)})
    if args.autotune:
        log(f"autotune: {s.autotune()}")


def ingest_workload(args):
    """Genome file -> resident genome: host readers vs device FASTA encoder and 2bit decoder on the headline genome."""
    total = int(args.genome_bp)
    lens = chrom_lengths(total, 24, args.seed)
    sg = cofpy.SynthGenome(args.seed, lens, n_frac=0.05, gap=(10_000, 1_000_000))
    with tempfile.TemporaryDirectory() as d:
        fa = os.path.join(d, "genome.fa")
        tb = os.path.join(d, "genome.2bit")
        t = time.time()
        sg.write_fasta(fa)
        size = os.path.getsize(fa)
        log(f"wrote {size / 1e9:.2f} GB FASTA in {time.time() - t:.1f} s")
        t = time.time()
        sg.write_2bit(tb)
        size2 = os.path.getsize(tb)
        log(f"wrote {size2 / 1e9:.2f} GB 2bit in {time.time() - t:.1f} s")
        res = {}
        for name, path, flags in (("host", fa, cofpy.COF_LOAD_HOST), ("device", fa, cofpy.COF_LOAD_DEVICE),
                                  ("2bit_host", tb, cofpy.COF_LOAD_HOST), ("2bit_device", tb, cofpy.COF_LOAD_DEVICE)):
            times = []
            for k in range(args.warmup + args.steps):
                s = cofpy.Search(["A" * 20 + "NGG"], ["N" * 20 + "NGG"], [0])
                t0 = time.perf_counter()
                s.load_file(path, flags)
                dt = time.perf_counter() - t0
                st = s.stats()
                s.close()
                if k >= args.warmup:
                    times.append(dt)
                log(f"{name} load: {dt:.3f} s (load_ms {st.load_ms:.1f}, device {st.load_device})")
            res[name] = round(sum(times) / len(times), 3)
    out = {"metric": "genome load (FASTA -> resident bit4 genome)", "value": res["device"], "unit": "s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(res["device"] * 1e3, 1),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u8 (bit4 encode)",
           "data": "synthetic",
           "config": {"workload": f"{total / 1e9:.2f} Gbp FASTA, 60-nt lines, 24 chromosomes, 5% N", "bytes": size},
           "host_reader_s": res["host"], "device_s": res["device"],
           "speedup": round(res["host"] / res["device"], 2),
           "twobit": {"bytes": size2, "host_reader_s": res["2bit_host"], "device_s": res["2bit_device"],
                      "speedup": round(res["2bit_host"] / res["2bit_device"], 2)}}
    print(json.dumps(out), flush=True)


def schedule_flags(args):
    return {"graph": 0, "eager": cofpy.COF_FLAG_NO_GRAPH, "sync": cofpy.COF_FLAG_SYNC_ROUNDS}[args.schedule]


def bulge_cpu_baseline(args, guides, plants):
    """CPU oracle on a bounded sample of the bulge workload (same genome model and query)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle
    sample = int(args.bulge_cpu_sample_bp)
    threads = min(16, os.cpu_count() or 1)
    sg = cofpy.SynthGenome(args.seed + 23, [sample], n_frac=0.05, gap=(10_000, 1_000_000),
                           repeats=(0.10, 20, 300, 0.10), plants=[(p[0], p[1], p[2], p[3], 1) for p in plants])
    with tempfile.TemporaryDirectory() as d:
        fa = os.path.join(d, "s.fa")
        sg.write_fasta(fa)
        in_text = f"{fa}\nNNNNNNNNNNNNNNNNNNNNNRG 2 2\n" + "".join(f"{g}NNN 4\n" for g in guides)
        n, st = oracle.run(in_text, os.path.join(d, "o.tsv"), batch_nucl=1 << 22, threads=threads)
    secs = st["search_seconds"]
    return {"value": round(len(guides) * sample / secs / 1e9, 4), "unit": "guide-Gbp/s", "cores": threads,
            "kind": "port", "sample": f"{len(guides)} guides x {sample / 1e6:.1f} Mbp repeat-rich synthetic, M4 D2 R2, "
                                      f"4 Mnt batches, search-only time {secs:.2f} s, {n} rows"}


def batched_guides(seed, n):
    """10,000-guide mix: len U{17..30}, PAM in {NRG, NGG, NNGRRT, TTTN}, M U{0..6}, 20% with D, R in U{1,2}."""
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        glen = rng.randint(17, 30)
        pam = rng.choice(["NRG", "NGG", "NNGRRT", "TTTN"])
        g = "".join(rng.choice("ACGT") for _ in range(glen))
        mm = rng.randint(0, 6)
        d, r = (rng.randint(1, 2), rng.randint(1, 2)) if rng.random() < 0.2 else (0, 0)
        if pam == "TTTN":
            out.append(("NNNN" + g, pam + "N" * glen, mm, d, r, pam + g, 4, glen))
        else:
            out.append((g + "N" * len(pam), "N" * glen + pam, mm, d, r, g + pam, 0, glen))
    return out


def batched_workload(args):
    """Batched search: all guides in one query vs one reference-format run per guide set."""
    total = int(args.genome_bp)
    lens = chrom_lengths(total, 24, args.seed)
    guides = batched_guides(args.seed + 11, args.batched_guides)
    plants = [(site, off, glen, min(mm, 3), 20) for (_, _, mm, _, _, site, off, glen) in guides[:2000]]
    t = time.time()
    sg = cofpy.SynthGenome(args.seed + 3, lens, n_frac=0.05, gap=(10_000, 1_000_000),
                           repeats=(0.10, 20, 300, 0.10), plants=plants)
    log(f"repeat-rich genome {total / 1e9:.3f} Gbp, {len(guides)} guides, generated in {time.time() - t:.1f} s")
    pats, filts, mms, ds, rs = zip(*[(g[0], g[1], g[2], g[3], g[4]) for g in guides])
    s = cofpy.Search(list(pats), list(filts), list(mms), list(ds), list(rs), count_only=True,
                     flags=schedule_flags(args))
    s.load_synth(sg)
    apply_tuning(s, args)
    res = {}
    # batched: one query
    times = []
    for k in range(0 if args.per_set_only else args.warmup + args.steps):
        t0 = time.perf_counter()
        n_plain = s.search()
        dt = time.perf_counter() - t0
        st = s.stats()
        if k >= args.warmup:
            times.append(dt)
        log(f"batched {'warmup' if k < args.warmup else 'step'}: {dt:.3f} s, plain hits {n_plain}, "
            f"bulge hits {st.bulge_hits}, candidates {st.n_candidates}, plain {st.kernel_ms_plain:.1f} ms, "
            f"bulge path {st.bulge_ms:.1f} ms ({st.n_pipeline_rounds} pipeline rounds)")
    if times:
        batched = sum(times) / len(times)
        res["batched"] = {"seconds": round(batched, 3), "plain_hits": n_plain, "bulge_hits": st.bulge_hits,
                          "bulge_checksum": st.bulge_checksum}
    if args.compare_plain_schedule or args.per_set_only:
        # plain schedule: one run per (filter, M, D, R) set
        sets = {}
        for g in guides:
            sets.setdefault((g[1], g[2], g[3], g[4]), []).append(g)
        t_sum, hp, hb, cs = 0.0, 0, 0, 0
        for si, (key, gs) in enumerate(sets.items()):
            if si % 100 == 0:
                log(f"plain schedule: run {si}/{len(sets)}, {t_sum:.2f} s so far")
            s.set_query([g[0] for g in gs], [key[0]] * len(gs), [key[1]] * len(gs), [key[2]] * len(gs),
                        [key[3]] * len(gs), count_only=True, flags=schedule_flags(args))
            t0 = time.perf_counter()
            hp += s.search()
            t_sum += time.perf_counter() - t0
            st = s.stats()
            hb += st.bulge_hits
            cs = (cs + st.bulge_checksum) % (1 << 64)
        log(f"plain schedule: {len(sets)} runs, {t_sum:.3f} s, plain hits {hp}, bulge hits {hb}")
        res["plain_schedule"] = {"seconds": round(t_sum, 3), "runs": len(sets), "plain_hits": hp, "bulge_hits": hb}
        res["plain_schedule"]["bulge_checksum"] = cs
        if times:
            res["speedup"] = round(t_sum / batched, 2)
            res["identical_counts"] = hp == res["batched"]["plain_hits"] and hb == res["batched"]["bulge_hits"]
            res["identical_checksum"] = cs == res["batched"]["bulge_checksum"]
        else:
            batched = t_sum
    st = s.stats()
    units = len(guides) * st.genome_nucl
    s.close()
    out = {"metric": "guide-Gbp/s at bit-exact hits (batched search)", "value": round(units / batched / 1e9, 3),
           "unit": "guide-Gbp/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(batched * 1e3, 1), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "u32 popcount / u32-u64 bit-parallel Myers", "data": "synthetic",
           "config": {"workload": f"batched: {total / 1e9:.3f} Gbp repeat-rich synthetic, {len(guides)} guides len 17-30, "
                                  "PAM {NRG,NGG,NNGRRT,TTTN}, M 0-6, 20% with D,R 1-2; bulge hits counted+checksummed",
                      "genome_bp": st.genome_nucl}}
    out.update(res)
    print(json.dumps(out), flush=True)


def tsv_workload(args):
    """Reference TSV output of the headline run with the most rows (NRG-20-M6 on the 3.1 Gbp genome):
    rows formatted on the device vs the host writer, into /dev/null (formatting) and into a file."""
    import hashlib
    total = int(args.genome_bp)
    lens = chrom_lengths(total, 24, args.seed)
    runs = make_queries(args.seed)
    plants = [p for r in runs for p in r["plants"]]
    sg = cofpy.SynthGenome(args.seed, lens, n_frac=0.05, gap=(10_000, 1_000_000), plants=plants)
    r = next(r for r in runs if r["name"] == "NRG-20-M6")
    s = cofpy.Search(r["pats"], [r["filter"]] * len(r["pats"]), [r["mm"]] * len(r["pats"]))
    s.load_synth(sg)
    rows = s.search()
    log(f"{r['name']}: {rows} rows")
    res = {}
    with tempfile.TemporaryDirectory(dir=args.tsv_dir) as d:
        digests = {}
        for mode in ("device", "host"):
            path = os.path.join(d, mode + ".tsv")
            s.write_tsv(path, host=mode == "host")
            digests[mode] = (os.path.getsize(path), hashlib.sha256(open(path, "rb").read()).hexdigest())
            os.remove(path)
            for target in ("/dev/null", path):
                times = []
                for k in range(args.warmup + args.steps):
                    s.write_tsv(target, host=mode == "host")
                    if k >= args.warmup:
                        times.append(s.stats().write_ms)
                    if target != "/dev/null":
                        os.remove(target)
                key = f"{mode}_{'devnull' if target == '/dev/null' else 'file'}_ms"
                res[key] = round(sum(times) / len(times), 2)
                log(f"{mode} -> {target}: {res[key]} ms")
    s.close()
    size, _ = digests["device"]
    out = {"metric": "TSV rows/s (reference row format, device formatter)",
           "value": round(rows / (res["device_devnull_ms"] / 1e3)), "unit": "rows/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": res["device_devnull_ms"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8 (text)",
           "data": "synthetic",
           "config": {"workload": f"TSV rows of NRG-20-M6 (100 guides, M6) on the {total / 1e9:.2f} Gbp headline genome, "
                                  "written to /dev/null; file figures beside", "genome_bp": total},
           "rows": rows, "bytes": size, "identical_bytes": digests["device"] == digests["host"],
           "host_rows_per_s": round(rows / (res["host_devnull_ms"] / 1e3)),
           "speedup_vs_host": round(res["host_devnull_ms"] / res["device_devnull_ms"], 2)}
    out.update(res)
    print(json.dumps(out), flush=True)


def cli_workload(args):
    """End to end as the reference CLI runs it (input file -> genome FASTA on disk -> TSV on disk): cof-cli on the
    GPU vs the oracle's CLI twin on 16 host threads, one headline run (NRG-20-M3, 100 guides) on the 3.1 Gbp
    genome, both with 4 Mnt batches so the row sets must be equal."""
    import subprocess
    total = int(args.genome_bp)
    lens = chrom_lengths(total, 24, args.seed)
    runs = make_queries(args.seed)
    plants = [p for r in runs for p in r["plants"]]
    r = next(r for r in runs if r["name"] == "NRG-20-M3")
    sg = cofpy.SynthGenome(args.seed, lens, n_frac=0.05, gap=(10_000, 1_000_000), plants=plants)
    cli = os.path.join(ROOT, "cas-offinder-rust-old_amd", "_build", "cof-cli")
    ref = os.path.join(ROOT, "oracle", "_build", "cof-ref")
    threads = min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory(dir=args.tsv_dir) as d:
        fa = os.path.join(d, "genome.fa")
        sg.write_fasta(fa)
        inp = os.path.join(d, "in.txt")
        with open(inp, "w") as f:
            f.write(f"{fa}\n{r['filter']}\n" + "".join(f"{p} {r['mm']}\n" for p in r["pats"]))
        env = dict(os.environ, COF_BATCH_NUCL=str(1 << 22))
        gpu_s = []
        for k in range(args.warmup + args.steps):
            t0 = time.perf_counter()
            subprocess.run([cli, inp, "G", os.path.join(d, "gpu.tsv")], env=env, check=True, capture_output=True)
            dt = time.perf_counter() - t0
            log(f"cof-cli: {dt:.2f} s")
            if k >= args.warmup:
                gpu_s.append(dt)
        t0 = time.perf_counter()
        subprocess.run([ref, inp, "C", os.path.join(d, "ref.tsv"), "--batch-nucl", str(1 << 22), "--threads",
                        str(threads)], check=True, capture_output=True)
        ref_s = time.perf_counter() - t0
        log(f"oracle CLI ({threads} threads): {ref_s:.2f} s")
        rows_gpu = sorted(open(os.path.join(d, "gpu.tsv")).read().splitlines()[1:])
        rows_ref = sorted(open(os.path.join(d, "ref.tsv")).read().splitlines()[1:])
    secs = sum(gpu_s) / len(gpu_s)
    out = {"metric": "end-to-end CLI wall time (input file to TSV, reference format)", "value": round(secs, 3),
           "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(secs * 1e3, 1),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u32 (bit4 / 2-bit planes)",
           "data": "synthetic",
           "config": {"workload": f"cof-cli on a {total / 1e9:.2f} Gbp FASTA, NRG-20-M3 (100 guides), 4 Mnt batches, "
                                  "process start to TSV written", "genome_bp": total},
           "rows": len(rows_gpu), "identical_rows": rows_gpu == rows_ref,
           "guide_gbp_s_end_to_end": round(len(r["pats"]) * total / secs / 1e9, 1),
           "cpu_baseline": {"value": round(ref_s, 2), "unit": "s", "cores": threads, "kind": "port",
                            "sample": "the same input through the oracle's CLI twin (cof-ref), full genome"}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--genome-bp", type=float, default=3.1e9)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--cpu-sample-bp", type=float, default=8e8)
    ap.add_argument("--bulge-cpu-sample-bp", type=float, default=5e5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--queue-depth", type=int, default=2,
                    help="plain workload: searches queued ahead of the one being finished (measured: depth 2 "
                         "fastest; 1 leaves the GPU idle behind a hit-heavy run's conversion, 3 and 6 put "
                         "more streams in flight than the process's hardware queues)")
    ap.add_argument("--no-overlap", action="store_true",
                    help="plain workload: finish each run before starting the next")
    ap.add_argument("--workload", choices=["plain", "bulge", "batched", "costmodel", "ingest", "tsv", "cli"],
                    default="plain")
    ap.add_argument("--tsv-dir", default=None, help="tsv: folder for the file-output timings (default: system temp)")
    ap.add_argument("--train-bp", type=float, default=1e8, help="costmodel: size of each training genome")
    ap.add_argument("--train-genomes", type=int, default=20)
    ap.add_argument("--train-steps", type=int, default=3000)
    ap.add_argument("--bulge-guides", type=int, default=100)
    ap.add_argument("--batched-guides", type=int, default=10_000)
    ap.add_argument("--compare-plain-schedule", action="store_true")
    ap.add_argument("--per-set-only", action="store_true",
                    help="batched workload: time only the one-run-per-set schedule (for runs too long to do both)")
    ap.add_argument("--schedule", choices=["graph", "eager", "sync"], default="graph",
                    help="bulge rounds: hipGraph pipeline, eager pipeline, or synchronous rounds")
    ap.add_argument("--autotune", action="store_true", help="run the autotuner on a 64 Mnt prefix first")
    ap.add_argument("--tune", action="append", default=[], metavar="FIELD=VALUE",
                    help="cof_tuning field for the bulge/batched/costmodel contexts (repeatable)")
    ap.add_argument("--no-verify", action="store_true",
                    help="plain workload: skip the check of every run's rows against the oracle digests")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="plain workload: run a 256 MiB device copy first (the traffic counters' calibration)")
    ap.add_argument("--manifest", default=None,
                    help="plain workload: write the runs in dispatch order (with --counters-from: read them)")
    ap.add_argument("--counters-from", nargs=4, default=None, metavar=("FETCH", "WRITE", "SQ", "MFMA"),
                    help="summarise four rocprofv3 --pmc folders of the plain workload into "
                         "profiles/plain_scan_counters.json (or, with --workload bulge, of the bulge workload "
                         "into profiles/bulge_counters.json)")
    ap.add_argument("--counters-out", default=None,
                    help="with --counters-from: where the JSON goes (default: the file under profiles/)")
    args = ap.parse_args()
    if args.counters_from:
        return summarize_bulge_counters(args) if args.workload == "bulge" else summarize_counters(args)
    if args.gpus != 1:
        print("bench.py measures one GPU only: the search is single-GPU by design "
              "(guide or genome sharding across GPUs needs only a host-side gather of hits)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "bulge":
        return bulge_workload(args)
    if args.workload == "batched":
        return batched_workload(args)
    if args.workload == "costmodel":
        return costmodel_workload(args)
    if args.workload == "ingest":
        return ingest_workload(args)
    if args.workload == "tsv":
        return tsv_workload(args)
    if args.workload == "cli":
        return cli_workload(args)

    copy_bytes = 256 << 20
    if args.calibrate_copy:
        calibration_copy()
    total = int(args.genome_bp)
    lens = chrom_lengths(total, 24, args.seed)
    runs = make_queries(args.seed)
    plants = [p for r in runs for p in r["plants"]]
    t = time.time()
    sg = cofpy.SynthGenome(args.seed, lens, n_frac=0.05, gap=(10_000, 1_000_000), plants=plants)
    log(f"genome: {total / 1e9:.2f} Gbp in {len(lens)} chromosomes, {len(plants)} plant sets, "
        f"generated in {time.time() - t:.1f} s")

    ctxs = []
    for r in runs:
        s = cofpy.Search(r["pats"], [r["filter"]] * len(r["pats"]), [r["mm"]] * len(r["pats"]))
        s.load_synth(sg)
        apply_tuning(s, args)
        ctxs.append(s)
    log(f"{len(ctxs)} contexts with the genome resident")

    kernel_ms = [0.0] * len(runs)
    search_ms = [0.0] * len(runs)
    convert_ms = [0.0] * len(runs)
    launches = [0] * len(runs)
    hits = [0] * len(runs)

    depth = max(1, args.queue_depth)

    def one_step(record):
        # runs i+1 .. i+depth are queued before run i's hits are copied back
        # and converted (cof_search_async / cof_search_wait; the scans still run
        # one after another on the GPU), unless --no-overlap
        def finish(i):
            hits[i] = ctxs[i].wait() if not args.no_overlap else ctxs[i].search()
            if record:
                st = ctxs[i].stats()
                kernel_ms[i] += st.kernel_ms_plain
                search_ms[i] += st.search_ms
                convert_ms[i] += st.convert_ms
                launches[i] = st.n_plain_launches
        for i, s in enumerate(ctxs):
            if not args.no_overlap:
                s.search_async()
                if i >= depth:
                    finish(i - depth)
            else:
                finish(i)
        if not args.no_overlap:
            for j in range(max(0, len(ctxs) - depth), len(ctxs)):
                finish(j)

    for w in range(args.warmup):
        one_step(False)
        log(f"warmup {w + 1}/{args.warmup} done")
    step_s = []
    for k in range(args.steps):
        t0 = time.perf_counter()
        one_step(True)
        step_s.append(time.perf_counter() - t0)
        log(f"step {k + 1}/{args.steps}: {step_s[-1] * 1e3:.1f} ms")

    genome_nucl = ctxs[0].stats().genome_nucl
    default_size = abs(genome_nucl - 3.1e9) < 1e6 and args.seed == 2024
    if args.manifest:
        with open(args.manifest, "w") as f:
            json.dump({"workload": "bench.py plain workload, 3.1 Gbp, 12 runs, --no-overlap", "copy_bytes": copy_bytes,
                       "runs": [[r["name"], launches[i], hits[i], r["L"], genome_nucl] for i, r in enumerate(runs)]}, f)
    # bit-exact check at full size: every run's sorted rows against the oracle's digests
    # (tests/golden/make_headline_digest.py), after the timed steps
    verify = None
    if not args.no_verify and default_size and os.path.exists(DIGESTS):
        with open(DIGESTS) as f:
            want = json.load(f)["runs"]
        verify = {}
        with tempfile.TemporaryDirectory() as d:
            for i, r in enumerate(runs):
                path = os.path.join(d, "rows.tsv")        # the hits of the last timed step
                ctxs[i].write_tsv(path)
                n, sha = rows_digest(path)
                os.remove(path)
                verify[r["name"]] = {"rows": n, "match": n == want[r["name"]]["rows"] and sha == want[r["name"]]["sha256"]}
        log(f"oracle digests: {sum(v['match'] for v in verify.values())}/{len(verify)} runs identical")
    units = sum(len(r["pats"]) for r in runs) * genome_nucl          # guide-bp per step
    step = sum(step_s) / len(step_s)
    value = units / step / 1e9
    kms = [km / args.steps for km in kernel_ms]
    alg_ops = sum(ops_per_guide_bp(r["L"]) * len(r["pats"]) * genome_nucl for r in runs)
    achieved = alg_ops / (sum(kms) / 1e3)
    # HBM roofline of the plain-scan kernel: algorithmic bytes of one launch =
    # the packed genome read once (0.5 B per position for all guides of the
    # launch) + 16 B per hit record; divided by the mean launch duration
    n_launch = sum(launches)
    alg_bytes = sum(launches[i] * genome_nucl / 2 + 16 * hits[i] for i in range(len(runs))) / n_launch
    launch_s = sum(kms) / 1e3 / n_launch
    hbm_achieved = alg_bytes / launch_s
    # HBM traffic per launch and issue rates: rocprofv3 --pmc figures of this
    # workload, read from the committed summary (bench.py --counters-from), and
    # null unless that summary was measured on the current plain.hip
    cdoc, csrc = measured_counters() if default_size else (None, "not the default workload")
    traffic = cdoc["traffic_bytes_per_launch"] if cdoc else None
    valu_issue = cdoc["valu_issue_frac"] if cdoc else None
    mfma_tops = cdoc["mfma_fp4_tops"] if cdoc else None
    mfma_busy = cdoc["mfma_busy_frac"] if cdoc else None
    per_run = [{"run": r["name"], "hits": hits[i], "kernel_ms": round(kms[i], 3),
                "oracle_rows_match": verify[r["name"]]["match"] if verify else None,
                "search_ms": round(search_ms[i] / args.steps, 3), "convert_ms": round(convert_ms[i] / args.steps, 3),
                "guide_gbp_s_kernel": round(len(r["pats"]) * genome_nucl / (kms[i] / 1e3) / 1e9, 1)}
               for i, r in enumerate(runs)]
    for s in ctxs:
        s.close()

    baseline = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        baseline = cpu_baseline(runs[0], int(args.cpu_sample_bp), args.seed, threads)
        log(f"cpu baseline: {baseline}")

    out = {
        "metric": "guide-Gbp/s at bit-exact hits (plain mismatch scan)",
        "value": round(value, 2),
        "unit": "guide-Gbp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp4 (matrix-core window scores: exact small integers in fp32) + u32 (bit4/plane ops)",
        "data": "synthetic",
        "config": {"workload": f"plain GPU mismatch scan: {genome_nucl / 1e9:.2f} Gbp synthetic genome (24 chroms, "
                               "41% GC, 5% N), 100 guides x len{20,25} x PAM{NRG,NNGRRT,TTTN} x M{3,6}, both "
                               "strands, 1000 planted near-hits per guide; 12 reference-format runs per step",
                   "genome_bp": genome_nucl, "guides_per_run": 100, "runs": len(runs),
                   "runs_overlapped": not args.no_overlap, "queue_depth": None if args.no_overlap else depth},
        "roofline": {"bound": "hbm", "achieved": round(hbm_achieved / 1e9, 1), "peak": round(HBM_PEAK / 1e9),
                     "unit": "GB/s", "frac": round(hbm_achieved / HBM_PEAK, 4), "traffic": traffic,
                     "kernel": "plain_scan_kernel", "launch_ms": round(launch_s * 1e3, 3),
                     "bytes_per_launch": round(alg_bytes),
                     "note": "the scan is bound by matrix-core and VALU issue, not by HBM: see counters"},
        "roofline_int32": {"bound": "int32-valu", "achieved": round(achieved / 1e12, 3),
                           "peak": round(INT32_PEAK / 1e12, 1), "unit": "TOP/s",
                           "frac": round(achieved / INT32_PEAK, 3),
                           "note": "achieved = brute-force int32 ops 2*(3*ceil(L/8)+1) per guide-bp / plain-scan "
                                   "kernel time; the kernel tests the PAM first and so executes fewer ops"},
        "counters": {"valu_issue_frac": valu_issue, "mfma_fp4_tops": mfma_tops,
                     "mfma_fp4_frac": round(mfma_tops * 1e12 / MFMA_FP4_PEAK, 3) if mfma_tops else None,
                     "mfma_busy_frac": mfma_busy,
                     "bank_conflict_cycles_per_lds": cdoc["bank_conflict_cycles_per_lds"] if cdoc else None,
                     "source": csrc,
                     "note": "measured issue of plain_scan_kernel (rocprofv3 --pmc, default workload only): VALU "
                             "instructions against the gfx950 issue peak; fp4 matrix-core ops against the dense "
                             "fp4 peak of 10.1 POP/s and matrix-core busy cycles"},
        "cpu_baseline": baseline,
        "bit_exact": (all(v["match"] for v in verify.values()) if verify is not None else None),
        "bit_exact_check": (f"sorted rows of all {len(verify)} runs vs the oracle's row counts and SHA-256 digests "
                            "(tests/golden/headline_3.1Gbp_digest.json)" if verify is not None else "not run"),
        "kernel_guide_gbp_s": round(units / (sum(kms) / 1e3) / 1e9, 1),
        "per_run": per_run,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
