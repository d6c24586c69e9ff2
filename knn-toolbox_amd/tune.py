"""Autotuner of the search planner: measures, for a grid of (precision, d, k, n, m), the planner
knobs that give the fastest search, and writes them as the tuning table the engine reads
(knn-toolbox_amd/tuning/gfx950.tsv; see knns_set_tuning_table in include/knns_native.h).

    python knn-toolbox_amd/tune.py [--quick] [--precisions bf16,fp32,fp8] [--merge] [--out PATH]

Knobs (the role of src/knnsplan.cu:43-57, which timed 80 full runs of two algorithms x four
stream counts; here every knob shapes the one fused kernel sequence):
  qb        query blocks of 32 per scan wave -- the tile shape: 32 qb queries x 32 corpus rows
            per MFMA tile, 4 query groups per workgroup sharing each corpus stage; 1 and 2 (two
            resident workgroups per CU) everywhere, 3 and 4 for FP8 rows up to 128 bytes and BF16
            rows of 64 dims and 3 for BF16 at d = 128 (two per CU), 4 (one per CU) at d = 128;
  min_waves slices per query group: the smallest count that fills that many waves of resident
            scan workgroups (2 to 8);
  ratio     corpus growth between threshold rounds (0 = x4 while the prefix is below m/16, then x2);
  sl_extra  shortlist offset from the precision's built-in rule (BF16, FP32: exactness does not
            depend on it -- the certificate re-searches what it cannot prove -- only the cost does;
            FP8 is not tuned, its shortlist sets its recall).
The knobs are searched coordinate by coordinate from the built-in setting (qb, then min_waves,
then ratio, then sl_extra); a change is kept only when it is faster by more than 2%. Every
candidate's output is checked identical to the first one's. Each point is the median device
time of --reps searches (corpus, queries and results resident, a 1024-centre Gaussian mixture,
sigma 0.1); corpus sizes are 1M rows, and 10M rows for the BF16 and FP8 points at d = 96, 128.
"""
from __future__ import annotations

import argparse
import datetime
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

# torch holds its own HIP runtime: it must be initialised before the library's is first loaded
import torch  # noqa: E402,F401

PKG = Path(__file__).resolve().parent
sys.path.insert(0, str(PKG))
import knns  # noqa: E402

QBS = (2, 1, 3, 4)             # the built-in choice for n > 32 first
WAVES = (5, 2, 3, 4, 6, 8)     # 5 = built-in
RATIOS = (0, 2, 3, 4)          # 0 = built-in schedule
EXTRAS = (0, -8, -16, -32, -48, 32)


def mixture(n: int, d: int, seed: int, centers: np.ndarray, chunk: int = 1 << 20) -> np.ndarray:
    g = np.random.Generator(np.random.PCG64(seed))
    out = np.empty((n, d), dtype=np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        a = g.integers(0, centers.shape[0], e - s)
        out[s:e] = centers[a] + 0.1 * g.standard_normal((e - s, d))
    return out


def set_table(lines: list[str]) -> None:
    fd, path = tempfile.mkstemp(suffix=".tsv")
    with os.fdopen(fd, "w") as f:
        f.write("\n".join(lines) + "\n")
    st = knns.native().knns_set_tuning_table(path.encode())
    os.unlink(path)
    if st:
        raise knns.KnnsError(st, knns.native().knns_last_error().decode())


def qb_ok(prec: str, d: int, qb: int) -> bool:
    """Tile shapes the scan is built for (others fall back to the built-in rule)."""
    if qb in (1, 2):
        return True
    if prec == "fp8":
        return qb in (3, 4) and d <= 128
    return prec == "bf16" and d in (64, 128)


class Timer:
    """Median device time of `reps` searches with queries and results resident in device memory."""

    def __init__(self, corpus, Q, k, precision, reps):
        import torch

        dev = torch.device("cuda:0")
        self.torch = torch
        self.corpus, self.n, self.k, self.prec, self.reps = corpus, Q.shape[0], k, precision, reps
        self.Qd = torch.from_numpy(np.ascontiguousarray(Q)).to(dev)
        self.Dd = torch.empty((self.n, k), dtype=torch.float32, device=dev)
        self.Id = torch.empty((self.n, k), dtype=torch.int32, device=dev)

    def __call__(self):
        run = lambda: self.corpus.search_device(self.Qd, self.n, self.k, self.Dd, self.Id, precision=self.prec)  # noqa: E731
        run()                                                  # warm (workspace, first launch)
        ts, st = [], None
        for _ in range(self.reps):
            run()
            st = knns.last_stats()
            ts.append(st["total_s"])
        ts.sort()
        return ts[len(ts) // 2], self.Id.cpu().numpy(), st


def tune_point(prec, d, k, n, m, timer, line_of):
    """Coordinate descent from the built-in setting; returns (knobs, ms, stats)."""
    cur = {"qb": 0, "waves": 5, "ratio": 0, "extra": 0}
    ref = None
    cache = {}

    def measure(knobs):
        key = tuple(sorted(knobs.items()))
        if key not in cache:
            set_table([line_of(knobs)])
            t, I, st = timer()
            nonlocal ref
            if ref is None:
                ref = I
            elif not np.array_equal(I, ref):
                raise RuntimeError(f"results changed with the knobs {knobs} at {prec} d={d} k={k} n={n} m={m}")
            cache[key] = (t, st)
        return cache[key]

    best_t, best_st = measure(cur)
    axes = [("qb", [q for q in QBS if qb_ok(prec, d, q)]), ("waves", WAVES), ("ratio", RATIOS)]
    if prec != "fp8":
        axes.append(("extra", EXTRAS))
    for name, values in axes:
        for v in values:
            cand = dict(cur, **{name: v})
            t, st = measure(cand)
            if t < best_t * 0.98:
                cur, best_t, best_st = cand, t, st
    return cur, best_t * 1e3, best_st


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--quick", action="store_true", help="one point (bf16, d=128, k=100, n=10000, m=1M)")
    ap.add_argument("--precisions", default="bf16,fp32,fp8", help="comma-separated: bf16, fp32, fp8")
    ap.add_argument("--no-10m", action="store_true", help="skip the 10M-row corpus points")
    ap.add_argument("--merge", action="store_true",
                    help="keep the entries of the other precisions from the table in the package")
    ap.add_argument("--out", default=str(PKG / "tuning" / "gfx950.tsv"))
    args = ap.parse_args()
    if knns.device_count() <= 0:
        print("no GPU visible", file=sys.stderr)
        return 3
    precs = ("bf16",) if args.quick else tuple(p.strip() for p in args.precisions.split(",") if p.strip())
    for p in precs:
        if p not in ("bf16", "fp32", "fp8"):
            raise SystemExit(f"unknown precision {p}")
    kept = []
    base = PKG / "tuning" / "gfx950.tsv"
    if args.merge and base.exists():
        for line in base.read_text().splitlines():
            f = line.split()
            if f and not line.startswith("#") and f[0] not in precs:
                kept.append(line)
    # (m, d, k values, n values) groups of points sharing one corpus
    if args.quick:
        groups = [(1_000_000, 128, (100,), (10000,))]
    else:
        groups = [(1_000_000, d, (10, 100), (256, 2048, 10000)) for d in (32, 64, 128, 256)]
        if not args.no_10m:
            groups += [(10_000_000, d, (10, 100), (10000,)) for d in (96, 128)]
    rows = []
    g = np.random.Generator(np.random.PCG64(77))
    for m, d, ks, ns in groups:
        centers = g.standard_normal((1024, d))
        R = mixture(m, d, 1, centers)
        Qall = mixture(max(ns), d, 2, centers)
        for prec in precs:
            if m > 1_000_000 and prec == "fp32":
                continue                      # the 10M points tune the low-precision shortlists
            with knns.Corpus(R, bf16=prec == "bf16", fp8=prec == "fp8") as corpus:
                for k in ks:
                    for n in ns:
                        Q = Qall[:n]
                        timer = Timer(corpus, Q, k, prec, args.reps)

                        def line_of(kn, prec=prec, d=d, k=k, n=n, m=m):
                            return f"{prec} {d} {k} {n} {m} {kn['waves']} {kn['ratio']} {kn['qb']} {kn['extra']}"
                        knobs, ms, st = tune_point(prec, d, k, n, m, timer, line_of)
                        rows.append((line_of(knobs), ms))
                        print(f"{prec} d={d} k={k} n={n} m={m}: qb={st['qb']} min_waves={knobs['waves']} "
                              f"ratio={knobs['ratio']} sl_extra={knobs['extra']} (K'={st['shortlist']}, "
                              f"re-searched {st['researched']}) {ms:.3f} ms", flush=True)
        del R
    set_table([])
    out = Path(args.out)
    out.parent.mkdir(parents=True, exist_ok=True)
    with open(out, "w") as f:
        f.write(f"# planner tuning table for gfx950 (MI355X), written by knn-toolbox_amd/tune.py\n"
                f"# {datetime.date.today().isoformat()}, 1024-centre Gaussian mixture corpora, median of "
                f"{args.reps} searches\n"
                f"# precision d k n m min_waves ratio qb sl_extra ms\n")
        for line in kept:
            f.write(line + "\n")
        for line, ms in rows:
            f.write(f"{line} {ms:.3f}\n")
    print(f"wrote {out} ({len(rows)} new entries, {len(kept)} kept)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
