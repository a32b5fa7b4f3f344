"""Combine twist-resolved n(mu) CSV files (the consumer of twist scans).

Same command line and output as the reference's
scripts/average_twist_outputs.py:8-116:

    python -m ftlm_extended_hubbard_development_amd.twist_average twist_*.csv \
        --mode observable-average|partition-sum [--sites L] [--single-beta B] \
        --output out.csv

Rows are grouped by (beta, mu). observable-average takes the plain mean of
n, charge_correlation and compressibility over twists and the log of the
mean partition function; partition-sum weights each twist by its partition
function, w_t = Z_t / sum Z (computed in log space), and rebuilds n, the
charge fluctuation and kappa from the mixed moments <N> and <N^2>.
"""
import argparse
import csv
import math
from collections import defaultdict

FIELDS = ["beta", "mu", "n", "charge_correlation", "compressibility", "twist_count", "mode", "log_partition"]


def log_sum_exp(values):
    top = max(values)
    return top + math.log(sum(math.exp(v - top) for v in values))


def read_rows(paths, single_beta=None):
    """Rows of the input files; single-beta files (no beta column) need
    single_beta (scripts/average_twist_outputs.py:15-38)."""
    for path in paths:
        with open(path, newline="") as f:
            reader = csv.DictReader(f)
            has_beta = "beta" in (reader.fieldnames or [])
            if not has_beta and single_beta is None:
                raise SystemExit(f"{path} has no beta column; pass --single-beta for legacy single-beta files")
            for row in reader:
                lp = row.get("log_partition")
                yield {
                    "beta": float(row["beta"]) if has_beta else float(single_beta),
                    "mu": float(row["mu"]),
                    "n": float(row["n"]),
                    "charge_correlation": float(row["charge_correlation"]),
                    "compressibility": float(row["compressibility"]),
                    "log_partition": float(lp) if lp not in (None, "") else None,
                }


def combine(rows, mode, sites=None):
    """One output record per (beta, mu), sorted; mode as on the command line
    (scripts/average_twist_outputs.py:59-111)."""
    if mode == "partition-sum" and (sites is None or sites <= 0):
        raise SystemExit("--sites is required and must be positive for --mode partition-sum")
    groups = defaultdict(list)
    for r in rows:
        groups[(r["beta"], r["mu"])].append(r)
    out = []
    for (beta, mu), grp in sorted(groups.items()):
        if mode == "observable-average":
            k = len(grp)
            n = sum(r["n"] for r in grp) / k
            cc = sum(r["charge_correlation"] for r in grp) / k
            kappa = sum(r["compressibility"] for r in grp) / k
            logs = [r["log_partition"] for r in grp if r["log_partition"] is not None]
            logz = log_sum_exp(logs) - math.log(len(logs)) if logs else ""
        elif mode == "partition-sum":
            if any(r["log_partition"] is None for r in grp):
                raise SystemExit("partition-sum mode requires log_partition in every input row")
            lz = [r["log_partition"] for r in grp]
            total = log_sum_exp(lz)
            wts = [math.exp(v - total) for v in lz]
            m1 = 0.0
            m2 = 0.0
            for w, r in zip(wts, grp):
                nt = r["n"] * sites
                n2t = r["charge_correlation"] * sites + nt * nt
                m1 += w * nt
                m2 += w * n2t
            n = m1 / sites
            cc = max(0.0, (m2 - m1 * m1) / sites)
            kappa = beta * cc
            logz = total
        else:
            raise SystemExit(f"unknown mode {mode}")
        out.append({"beta": beta, "mu": mu, "n": n, "charge_correlation": cc, "compressibility": kappa,
                    "twist_count": len(grp), "mode": mode, "log_partition": logz})
    return out


def write(path, records):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for rec in records:
            w.writerow(rec)


def main(argv=None):
    p = argparse.ArgumentParser(description="Average or partition-sum FTLM n(mu) CSV files over twists.")
    p.add_argument("inputs", nargs="+", help="twist-resolved CSV files")
    p.add_argument("--output", required=True, help="output CSV path")
    p.add_argument("--mode", choices=["observable-average", "partition-sum"], default="observable-average",
                   help="average observables equally, or combine twists by summing partition functions")
    p.add_argument("--sites", type=int, help="number of lattice sites; required for partition-sum")
    p.add_argument("--single-beta", type=float,
                   help="beta value for legacy single-beta files without a beta column")
    a = p.parse_args(argv)
    if a.mode == "partition-sum" and (a.sites is None or a.sites <= 0):
        raise SystemExit("--sites is required and must be positive for --mode partition-sum")
    write(a.output, combine(read_rows(a.inputs, a.single_beta), a.mode, a.sites))


if __name__ == "__main__":
    main()
