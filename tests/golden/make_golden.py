"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in the build container, where the reference sources are mounted:

    make -C oracle ref            # builds oracle/_ref/ref_dump from /root/reference
    python tests/golden/make_golden.py

For every case it runs oracle/_ref/ref_dump (the reference's own ddCOSMO
module driven through main.f90's calling sequence) and stores the results in
full FP64 as <case>.npz next to the input deck it used.  The six molecules of
the reference's test harness (tests/*.txt and their golden logs) are copied
as data; the synthetic configurations are written by ddpcm_amd.inputs.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ddpcm_amd import inputs  # noqa: E402

REF = os.environ.get("DDPCM_REFERENCE", "/root/reference")
DUMP = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
MOLECULES = ["formaldehyde", "benzene", "caffeine", "taxol", "ala25", "crambin"]


def read_dump(prefix):
    hdr = {}
    for line in open(prefix + ".hdr"):
        t = line.split()
        if t[0] == "arrays":
            continue
        hdr[t[0]] = float(t[1]) if any(c in t[1] for c in ".E") else int(t[1])
    n, ng, ny, nc, nnl = hdr["nsph"], hdr["ngrid"], hdr["nylm"], hdr["ncav"], hdr["nnl"]
    raw = open(prefix + ".bin", "rb").read()
    off = [0]

    def take(cnt, shape, dt=np.float64):
        a = np.frombuffer(raw, dtype=dt, count=cnt, offset=off[0])
        off[0] += cnt * np.dtype(dt).itemsize
        return a.reshape(shape).copy()

    d = {}
    d["grid"] = take(3 * ng, (ng, 3)); d["w"] = take(ng, (ng,))
    d["basis"] = take(ny * ng, (ng, ny)); d["facs"] = take(ny, (ny,)); d["facl"] = take(ny, (ny,))
    d["inl"] = take(n + 1, (n + 1,), np.int32); d["nl"] = take(nnl, (nnl,), np.int32)
    d["fi"] = take(ng * n, (n, ng)); d["ui"] = take(ng * n, (n, ng))
    if hdr["igrad"]:
        d["zi"] = take(3 * ng * n, (n, ng, 3))
    d["ccav"] = take(3 * nc, (nc, 3)); d["phi"] = take(nc, (nc,))
    d["psi"] = take(ny * n, (n, ny)); d["rhs"] = take(ny * n, (n, ny))
    d["sigma"] = take(ny * n, (n, ny))
    if hdr["igrad"]:
        d["s"] = take(ny * n, (n, ny)); d["xi"] = take(ng * n, (n, ng))
        d["fxdd"] = take(3 * n, (n, 3)); d["zeta"] = take(nc, (nc,)); d["fx"] = take(3 * n, (n, 3))
    if off[0] < len(raw):
        d["probe_v"] = take(ny * n, (n, ny)); d["probe_lx"] = take(ny * n, (n, ny))
        d["probe_lstarx"] = take(ny * n, (n, ny))
    return hdr, d


def run_case(name, deck, keep):
    prefix = os.path.join("/tmp", "golden_" + name)
    out = subprocess.run([DUMP, deck, prefix, "probe"], capture_output=True, text=True,
                         check=True).stdout
    iters = [line for line in out.splitlines() if "iter=" in line]
    hdr, d = read_dump(prefix)
    # iteration counts: the residual lines restart at 1 for the adjoint solve
    counts, cur = [], 0
    for line in iters:
        it = int(line.split("iter=")[1].split()[0])
        if it == 1 and cur:
            counts.append(cur)
        cur = it
    counts.append(cur)
    norms = [float(line.split(":")[-1].replace("D", "E")) for line in iters]
    rec = {"esolv": np.float64(hdr["esolv"]), "ncav": np.int64(hdr["ncav"]),
           "niter_direct": np.int64(counts[0]),
           "niter_adjoint": np.int64(counts[1] if len(counts) > 1 else 0),
           "residuals": np.array(norms), "hnorm_probe": np.float64(hdr["hnorm_probe"])}
    for k in keep:
        if k in d:
            rec[k] = d[k]
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **rec)
    print("%-14s n=%5d ncav=%7d E=%.12e iters=%s" % (name, hdr["nsph"], hdr["ncav"],
                                                    hdr["esolv"], counts))


FULL = ["grid", "w", "basis", "facs", "facl", "inl", "nl", "fi", "ui", "zi", "ccav", "phi",
        "psi", "rhs", "sigma", "s", "xi", "fxdd", "zeta", "fx", "probe_v", "probe_lx",
        "probe_lstarx"]
SOLUTION = ["phi", "rhs", "sigma", "s", "fxdd", "zeta", "fx", "probe_v", "probe_lx",
            "probe_lstarx", "inl", "nl"]


def main():
    if not os.path.exists(DUMP):
        sys.exit("build the reference first: make -C oracle ref")
    os.makedirs(os.path.join(HERE, "inputs"), exist_ok=True)
    os.makedirs(os.path.join(HERE, "logs"), exist_ok=True)
    only = sys.argv[1:]   # case names to (re)write; all when empty
    for m in MOLECULES:
        if only and m not in only:
            continue
        deck = os.path.join(HERE, "inputs", m + ".txt")
        shutil.copy(os.path.join(REF, "tests", m + ".txt"), deck)
        shutil.copy(os.path.join(REF, "tests", m + ".log"), os.path.join(HERE, "logs", m + ".log"))
        keep = FULL if m in ("formaldehyde", "benzene", "caffeine") else SOLUTION
        run_case(m, deck, keep)
    synth = {
        "c1": (inputs.config_input("c1"), FULL),
        "c1_forces": (inputs.config_input("c1", igrad=1), FULL),
        "c2_small": (inputs.config_input("c2", n=300, seed=2300), SOLUTION),
        "c2": (inputs.config_input("c2"), ["sigma", "s", "fx", "fxdd"]),
        "ion": (inputs.DDInput(0, 1, 6, 302, 10, 1, 78.39, 0.1, np.array([1.0]),
                               np.zeros((1, 3)), np.array([1.5])), FULL),
        # the large workload's lmax and grid on a small system, and higher lmax
        "c3_small": (inputs.config_input("c3", n=200, seed=20200), SOLUTION),
        "l12": (inputs.config_input("c1", n=40, seed=4012, lmax=12, ngrid=590, igrad=1),
                SOLUTION),
        "l16": (inputs.config_input("c1", n=12, seed=4016, lmax=16, ngrid=1202, igrad=1),
                SOLUTION),
    }
    for name, (d, keep) in synth.items():
        if only and name not in only:
            continue
        deck = os.path.join(HERE, "inputs", name + ".txt")
        inputs.write_input(deck, d)
        run_case(name, deck, keep)


if __name__ == "__main__":
    main()
