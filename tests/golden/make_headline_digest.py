"""Oracle digests of the headline workload (bench.py's default plain workload), for the at-size parity check.

Builds bench.py's seeded 3.1 Gbp synthetic genome (24 chromosomes, 41% GC, 5%
N in 10 kbp-1 Mbp gaps, 1,000 planted near-hits per guide), writes it as FASTA
into a scratch folder and runs the CPU oracle (oracle/cof_ref.cpp, the
restatement of the reference CPU path) on each of the 12 reference-format
runs (100 guides x len {20, 25} x PAM {NRG, NNGRRT, TTTN} x M {3, 6}) with
the batch size the benchmark's contexts use (1 Gnt, the reference GPU path's
batch).  For every run it records the row count and the SHA-256 of the sorted
data rows (each followed by a newline, header excluded) into
headline_3.1Gbp_digest.json beside this script; bench.py formats the GPU rows
of every run after its timed steps and compares them with these digests.

Runs on the CPU only (about half an hour on 8 cores, ~3.2 GB of scratch
disk): python tests/golden/make_headline_digest.py [--scratch DIR] [--jobs J]
"""
import argparse
import concurrent.futures as cf
import hashlib
import json
import os
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import bench  # noqa: E402  (workload definition: chromosome lengths, runs, plants)
import cofpy  # noqa: E402
import oracle  # noqa: E402

OUT = os.path.join(HERE, "headline_3.1Gbp_digest.json")


def rows_digest(path):
    """(row count, SHA-256 of the sorted data rows) of a reference-format TSV."""
    rows = cofpy.read_rows(path)
    h = hashlib.sha256()
    for r in rows:
        h.update(r.encode())
        h.update(b"\n")
    return len(rows), h.hexdigest()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scratch", default=None, help="folder for the FASTA and the row files (default: system temp)")
    ap.add_argument("--jobs", type=int, default=2, help="oracle runs at a time")
    ap.add_argument("--threads", type=int, default=4, help="oracle threads per run (a 1 Gnt batch is one task)")
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--genome-bp", type=float, default=3.1e9)
    args = ap.parse_args()
    total = int(args.genome_bp)
    lens = bench.chrom_lengths(total, 24, args.seed)
    runs = bench.make_queries(args.seed)
    plants = [p for r in runs for p in r["plants"]]
    with tempfile.TemporaryDirectory(dir=args.scratch) as d:
        t = time.time()
        sg = cofpy.SynthGenome(args.seed, lens, n_frac=0.05, gap=(10_000, 1_000_000), plants=plants)
        fa = os.path.join(d, "headline.fa")
        sg.write_fasta(fa)
        del sg
        print(f"genome written in {time.time() - t:.0f} s", flush=True)

        def one(r):
            t0 = time.time()
            in_text = f"{fa}\n{r['filter']}\n" + "".join(f"{p} {r['mm']}\n" for p in r["pats"])
            out = os.path.join(d, r["name"] + ".tsv")
            oracle.run(in_text, out, batch_nucl=1 << 30, threads=args.threads)
            n, sha = rows_digest(out)
            os.remove(out)
            print(f"{r['name']}: {n} rows in {time.time() - t0:.0f} s", flush=True)
            return r["name"], {"rows": n, "sha256": sha}

        with cf.ThreadPoolExecutor(args.jobs) as ex:
            res = dict(ex.map(one, runs))
    doc = {"workload": "bench.py default plain workload", "seed": args.seed, "genome_bp": total,
           "batch_nucl": 1 << 30, "digest": "sha256 of the sorted data rows, each followed by a newline",
           "runs": {r["name"]: res[r["name"]] for r in runs}}
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1)
        f.write("\n")
    print(f"wrote {OUT}", flush=True)


if __name__ == "__main__":
    main()
