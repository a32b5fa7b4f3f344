"""Combine twist-resolved n(mu) CSV files: the consumer of a twist scan
(--twist-grid / --twist-list runs write one CSV per twist).

Command line, messages and output columns are those of the reference's
scripts/average_twist_outputs.py:40-116 (the drop-in contract):

    python -m ftlm_extended_hubbard_development_amd.twist_average twist_*.csv \
        --mode observable-average|partition-sum [--sites L] [--single-beta B] \
        --output out.csv

The implementation is a columnar one: every input row becomes one entry of
flat numpy columns (beta, mu, n, charge_correlation, compressibility,
log Z; NaN where a file has no log_partition), rows are grouped by
(beta, mu) through np.unique, and each mode is a handful of segmented sums
over the groups:

* observable-average: n, charge_correlation and compressibility are the
  plain means over the twists of a group; log_partition is the log of the
  mean partition function, log(sum_t Z_t / T), over the rows that carry one
  (empty when none does).
* partition-sum: the twists form one ensemble, Z = sum_t Z_t, so twist t
  enters with weight Z_t / Z (computed from log Z_t relative to the group's
  largest); <N> and <N^2> are the weighted means of N_t = n_t L and
  N2_t = cc_t L + N_t^2, n = <N> / L, cc = max(0, <N^2> - <N>^2) / L,
  kappa = beta cc, log_partition = log Z.
"""
import argparse
import csv

import numpy as np

FIELDS = ["beta", "mu", "n", "charge_correlation", "compressibility", "twist_count", "mode", "log_partition"]
MODES = ("observable-average", "partition-sum")


class Table:
    """Flat columns of all input rows, in input order."""

    COLS = ("beta", "mu", "n", "charge_correlation", "compressibility", "log_partition")

    def __init__(self):
        self.cols = {c: [] for c in self.COLS}

    def add_file(self, path, single_beta):
        with open(path, newline="") as f:
            reader = csv.DictReader(f)
            names = reader.fieldnames or []
            if "beta" not in names and single_beta is None:
                raise SystemExit(f"{path} has no beta column; pass --single-beta for legacy single-beta files")
            for row in reader:
                self.cols["beta"].append(float(row["beta"]) if "beta" in names else float(single_beta))
                for c in ("mu", "n", "charge_correlation", "compressibility"):
                    self.cols[c].append(float(row[c]))
                lz = row.get("log_partition")
                self.cols["log_partition"].append(float(lz) if lz not in (None, "") else np.nan)

    def arrays(self):
        return {c: np.asarray(v, dtype=np.float64) for c, v in self.cols.items()}


def load(paths, single_beta=None):
    t = Table()
    for p in paths:
        t.add_file(p, single_beta)
    return t.arrays()


def _groups(a):
    """(beta, mu) keys in ascending order, each row's group and the counts."""
    keys, inv, cnt = np.unique(np.stack([a["beta"], a["mu"]], axis=1), axis=0, return_inverse=True,
                               return_counts=True)
    return keys, inv.ravel(), cnt


def _segment_sum(values, inv, ngroups):
    return np.bincount(inv, weights=values, minlength=ngroups)


def _segment_max(values, inv, ngroups):
    out = np.full(ngroups, -np.inf)
    np.maximum.at(out, inv, values)
    return out


def combine(a, mode, sites=None):
    """Records (dicts with FIELDS) per (beta, mu) in ascending order."""
    if mode not in MODES:
        raise SystemExit(f"unknown mode {mode}")
    if mode == "partition-sum" and (sites is None or sites <= 0):
        raise SystemExit("--sites is required and must be positive for --mode partition-sum")
    keys, inv, cnt = _groups(a)
    G = len(keys)
    lz = a["log_partition"]
    have = ~np.isnan(lz)
    if mode == "partition-sum" and not have.all():
        raise SystemExit("partition-sum mode requires log_partition in every input row")
    # log sum_t Z_t over the rows carrying log Z, relative to the group's top
    top = _segment_max(np.where(have, lz, -np.inf), inv, G)
    nz = _segment_sum(have.astype(np.float64), inv, G)
    safe_top = np.where(np.isfinite(top), top, 0.0)
    rel = np.where(have, np.exp(np.where(have, lz, 0.0) - safe_top[inv]), 0.0)
    log_total = safe_top + np.log(np.where(nz > 0, _segment_sum(rel, inv, G), 1.0))
    if mode == "observable-average":
        n = _segment_sum(a["n"], inv, G) / cnt
        cc = _segment_sum(a["charge_correlation"], inv, G) / cnt
        kappa = _segment_sum(a["compressibility"], inv, G) / cnt
        logz = [float(v) if k > 0 else "" for v, k in zip(log_total - np.log(np.maximum(nz, 1.0)), nz)]
    else:
        L = float(sites)
        w = np.exp(lz - log_total[inv])
        N1 = a["n"] * L
        N2 = a["charge_correlation"] * L + N1 * N1
        m1 = _segment_sum(w * N1, inv, G)
        m2 = _segment_sum(w * N2, inv, G)
        n = m1 / L
        cc = np.maximum(0.0, (m2 - m1 * m1) / L)
        kappa = keys[:, 0] * cc
        logz = [float(v) for v in log_total]
    return [{"beta": float(keys[g, 0]), "mu": float(keys[g, 1]), "n": float(n[g]), "charge_correlation": float(cc[g]),
             "compressibility": float(kappa[g]), "twist_count": int(cnt[g]), "mode": mode, "log_partition": logz[g]}
            for g in range(G)]


def write(path, records):
    with open(path, "w", newline="") as f:
        out = csv.writer(f)
        out.writerow(FIELDS)
        for r in records:
            out.writerow([r[k] for k in FIELDS])


def parser():
    p = argparse.ArgumentParser(description="Average or partition-sum FTLM n(mu) CSV files over twists.")
    p.add_argument("inputs", nargs="+", help="twist-resolved CSV files")
    p.add_argument("--output", required=True, help="output CSV path")
    p.add_argument("--mode", choices=list(MODES), default="observable-average",
                   help="average observables equally, or combine twists by summing partition functions")
    p.add_argument("--sites", type=int, help="number of lattice sites; required for partition-sum")
    p.add_argument("--single-beta", type=float, help="beta value for legacy single-beta files without a beta column")
    return p


def main(argv=None):
    args = parser().parse_args(argv)
    if args.mode == "partition-sum" and (args.sites is None or args.sites <= 0):
        raise SystemExit("--sites is required and must be positive for --mode partition-sum")
    write(args.output, combine(load(args.inputs, args.single_beta), args.mode, args.sites))


if __name__ == "__main__":
    main()
