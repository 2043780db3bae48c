"""Extract the reference's published output checksums into a fixture.

Reads the benchmark CSVs that ship with the reference (bench/*.csv and
bench/compile-20261011-followup/*.csv; data files, not code) and writes
reference_checksums.json: {modulus: {op: {size: fnv1a_hex}}}.  Each row's
checksum is the FNV-1a hash of the full output limbs (bench/bench.cpp:152-159)
for the deterministic inputs of bench/bench.cpp:216-373.

Modulus of a file: its `modulus` column when present, else the build the
file name refers to (bench/compile-20261011-followup/benchmark-builds.json:
crt-one -> 3, crt-two -> 100003, binary -> 2, full -> 1224736769,
crt -> 1000000007), else the bench default 998244353.  A
(modulus, op, size) seen with two different checksums is dropped and listed.

Run in a container that has the reference checkout:
    python tests/golden/make_reference_checksums.py /root/reference
"""
import csv
import glob
import json
import os
import sys

KEEP = {"conv", "conv-trunc", "conv-tail", "square", "skinny", "inv", "log", "exp", "sqrt",
        "pow", "ntt-forward", "ntt-inverse", "inv-short", "inv-linear", "exp-linear",
        "div", "mod", "divmod"}


def file_modulus(name):
    for tag, mod in (("crt-one", "3"), ("crt-two", "100003"), ("binary", "2"),
                     ("full", "1224736769"), ("crt", "1000000007")):
        if tag in name:
            return mod
    return "998244353"


def main(root):
    files = sorted(glob.glob(os.path.join(root, "bench", "*.csv")) +
                   glob.glob(os.path.join(root, "bench", "compile-20261011-followup", "*.csv")))
    table, conflicts, sources = {}, set(), {}
    for path in files:
        name = os.path.basename(path)
        with open(path, newline="") as fh:
            rows = list(csv.DictReader(fh))
        if not rows or "checksum" not in rows[0] or "operation" not in rows[0]:
            continue
        for row in rows:
            if "modulus" in row and row["modulus"]:
                mod = row["modulus"]
            else:
                mod = file_modulus(name)
            op, size, ck = row["operation"], row.get("size", ""), row["checksum"].strip().lower()
            if op not in KEEP or not size.isdigit() or len(ck) != 16:
                continue
            key = (mod, op, size)
            old = table.get(key)
            if old is not None and old != ck:
                conflicts.add(key)
            table[key] = ck
            sources.setdefault(key, os.path.relpath(path, root))
    out = {}
    for (mod, op, size), ck in sorted(table.items()):
        if (mod, op, size) in conflicts:
            continue
        out.setdefault(mod, {}).setdefault(op, {})[size] = ck
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reference_checksums.json"), "w") as fh:
        json.dump({"conflicts_dropped": sorted("/".join(k) for k in conflicts), "checksums": out},
                  fh, indent=1, sort_keys=True)
    n = sum(len(v2) for v in out.values() for v2 in v.values())
    print(f"{n} checksums from {len(files)} files; {len(conflicts)} conflicting keys dropped")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
