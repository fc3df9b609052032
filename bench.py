"""Energy + forces ddCOSMO solves per second on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload c4|c2|c3|c5]

The default workload (c4) is the metric's own comparison on the largest
single-GPU configuration: 512 synthetic molecules of 20..800 atoms (113 185
spheres; lmax 6, 302 Lebedev points, eps 78.39, eta 0.1, conv 1e-8, energy +
forces) on one batched handle, one step being the whole batch from
coordinates already resident in HBM: device setup (neighbours, fi/ui/zi,
cavity, triple lists), RHS (solute potential at the cavity points,
projection), direct and adjoint Jacobi/DIIS solves with energies, analytic
forces with zeta and both field terms.

  value      solves/s of the mixed precision (FP32 sweeps and products on
             the increments, the residual carried in FP64 and refreshed by
             the FP64 operator every eighth round) with the learned initial
             guess, its
             network trained beforehand on other molecules of the same sizes
             (training batches drawn with other seeds, untimed)
  fp64       the plain FP64 path on the same handle (the reference's
             arithmetic and iteration counts), and speedup_vs_fp64 = their
             ratio; per-stage times and iterations of both
  roofline   the dominant kernel of the timed (mixed) solves
  cpu_baseline  the reference's own OpenMP Fortran program (oracle/_ref/
             main.exe, built from its sources) on the first molecules of the
             batch, one process each, for about 20 s

c2 is the 2000-atom system (lmax 8, 590 points), c3 the 20000-atom one
(lmax 10, 1202 points): an FP64 line with the mixed precision beside it.
c5 is the 1000-atom trajectory timed under each initial-guess mode.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "E+forces solves/s at conv 1e-8: mixed+learned vs plain FP64 GPU; HBM GB/s"
HBM_PEAK_GBS = 8000.0        # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6      # MI355X FP64, vector and dense matrix alike
FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 dense matrix (vector FP32 reaches it only packed)
KERNELS = ("sweep_l", "sweep_lstar", "project", "expand", "jacobi", "diis", "extrap", "field",
           "potential", "forces", "sweep_pair")


def diis_depths(niter, ndiis=25):
    """(nmat after the drop-10 rule) per iteration of jacobi_diis.f90's DIIS:
    nmat runs 1..25, then 16..25, ... (jacobi_diis.f90:216-248)."""
    out, nmat = [], 1
    for _ in range(niter):
        if nmat >= ndiis:
            nmat -= 10
        out.append(nmat)
        nmat += 1
    return out


def algorithmic_work(h, st, iters=None):
    """Per kernel class: (unit, work per solve, launches per solve) for an
    energy + forces solve.  Counting convention: sweeps T (7 nylm + 40) flops
    per operator apply; MFMA products 2 nylm ngrid N flops; all-pairs
    20 ncav N flops per pass (the potential pass also returns the field, the
    field kernel is the zeta field at the nuclei); Jacobi update (4 + m) n 8
    bytes (rhs, y, x, m - 1 history residuals read; x_new, e written) and
    extrapolation (m + 2) n 8 bytes at DIIS depth m.  A batch counts each
    molecule's own vector length, iterations and all-pairs sums (iters:
    per-molecule (sizes, ncav, direct, adjoint)); its sweeps and products run
    over the whole batch until the last molecule converges."""
    T = st.ntriples
    sweep = T * (7.0 * h.nylm + 40.0)
    mm = 2.0 * h.nylm * h.ngrid * h.nsph
    if iters is None:
        iters = [(h.nsph, h.ncav, st.niter_direct, st.niter_adjoint)]
    jac = ext = pairs = 0.0
    for nq, cq, ndq, naq in iters:
        n = h.nylm * nq
        depth = diis_depths(ndq) + diis_depths(naq)
        jac += sum((4 + m) * n * 8.0 for m in depth)
        ext += sum((m + 2) * n * 8.0 for m in depth)
        pairs += 20.0 * cq * nq
    nd = max(v[2] for v in iters)
    na = max(v[3] for v in iters)
    # with forces the two solves iterate in lockstep: one sweep, Jacobi, DIIS
    # and extrapolation launch per round serves both
    rounds = max(nd, na)
    return {
        "sweep_pair": ("flop", sweep * (nd + na), rounds),
        "project": ("flop", mm * (nd + 1), nd + 1),
        "expand": ("flop", mm * (na + 2), na + 2),
        "jacobi": ("byte", jac, rounds),
        "extrap": ("byte", ext, rounds),
        "field": ("flop", pairs, 1),
        "potential": ("flop", pairs, 1),
    }


def kernel_table(h, iters_fn=None):
    """One solve with every kernel class timed: per-class launches, time and
    achieved rate against its roofline."""
    for k in KERNELS:
        h.kernel_timer(k, True)
    h.run()
    work = algorithmic_work(h, h.stats(), iters_fn() if iters_fn else None)
    table = {}
    for k in KERNELS:
        n, ms = h.kernel_timer_read(k)
        h.kernel_timer(k, False)
        row = {"launches": n, "ms": ms}
        if k in work and ms > 0:
            unit, w, _ = work[k]
            if unit == "flop":
                row["tflops"] = w / (ms * 1e-3) / 1e12
                row["frac_fp64_peak"] = row["tflops"] / FP64_PEAK_TFLOPS
            else:
                row["gbs"] = w / (ms * 1e-3) / 1e9
                row["frac_hbm_peak"] = row["gbs"] / HBM_PEAK_GBS
        table[k] = row
    return table


def _reference_wall(d, cores, timeout):
    """Process wall time of one solve of deck d by the reference program."""
    exe = os.path.join(ROOT, "oracle", "_ref", "main.exe")
    if not os.path.exists(exe):
        return None, None
    from ddpcm_amd import inputs
    tmp = tempfile.mkdtemp(prefix="ddpcm_ref_")
    try:
        dd = inputs.DDInput(1, cores, d.lmax, d.ngrid, d.iconv, d.igrad, d.eps, d.eta,
                            d.charge, d.xyz, d.rvdw)
        inputs.write_input(os.path.join(tmp, "Input.txt"), dd)
        env = dict(os.environ, OMP_NUM_THREADS=str(cores))
        t0 = time.perf_counter()
        r = subprocess.run([exe], cwd=tmp, capture_output=True, text=True, timeout=timeout,
                           env=env)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            return None, None
        m = re.search(r"solvation energy \(kcal/mol\):\s*([-0-9.Ee+]+)", r.stdout)
        return wall, (float(m.group(1)) if m else None)
    except subprocess.TimeoutExpired:
        return None, None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline(d, cores, timeout=300):
    """One solve of the workload by the reference program on `cores` threads."""
    wall, e_kcal = _reference_wall(d, cores, timeout)
    if wall is None:
        return None
    return {"value": 1.0 / wall, "unit": "solves/s", "cores": cores, "kind": "reference",
            "sample": "1 E+F solve of the %d-atom workload by the reference program "
                      "(oracle/_ref/main.exe, flang -O3 -fopenmp, nproc=%d), process wall "
                      "time %.2f s" % (d.n, cores, wall),
            "energy_kcal": e_kcal}


def cpu_baseline_batch(decks, cores, budget_s=20.0):
    """The reference program on the batch's molecules one after another (it
    holds one system per process), for about budget_s seconds."""
    tot, done = 0.0, 0
    for d in decks:
        wall, _ = _reference_wall(d, cores, 120)
        if wall is None:
            return None
        tot += wall
        done += 1
        if tot >= budget_s:
            break
    return {"value": done / tot, "unit": "solves/s", "cores": cores, "kind": "reference",
            "sample": "the first %d molecules of the batch (%d atoms in all), one reference "
                      "process each (oracle/_ref/main.exe, flang -O3 -fopenmp, nproc=%d), "
                      "%.2f s of process wall time" % (done, sum(x.n for x in decks[:done]),
                                                       cores, tot)}


def copy_calibration(nbytes=1 << 30):
    """One device-to-device copy of nbytes from a source that was evicted
    from the caches (a second buffer of the same size is written after it)."""
    import torch
    n = nbytes // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    evict = torch.full((n,), 2.0, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "check": float(dst[-1].item() + evict[0].item())}))


def traffic_from_profiles(kernel, tag):
    """HBM bytes per launch of `kernel` from profiles/<kernel>_<tag>_traffic.md
    (FETCH_SIZE + WRITE_SIZE counters scaled by the copy calibration), if
    that summary exists for this workload."""
    import glob
    for path in glob.glob(os.path.join(ROOT, "profiles", "%s_%s_traffic.md" % (kernel, tag))):
        with open(path) as f:
            for line in f:
                m = re.match(r"\s*traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", line)
                if m:
                    return float(m.group(1))
    return None


def trajectory_bench(args):
    """c5: one 1000-atom system (lmax 8, 590 points) along a trajectory of
    Gaussian steps (0.05 angstrom per atom and frame).  Each guess mode runs
    the same frames; the learned modes train on the first `--train` frames'
    solutions (8 Adam steps per frame) and all modes are timed on the frames
    after them; the last mode adds the mixed precision to the previous +
    learned guess.  value: solves/s of the learned guess."""
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import (DDCosmo, DDX_GUESS_LEARNED, DDX_GUESS_LEARNED_PREVIOUS,
                               DDX_GUESS_PREVIOUS, DDX_GUESS_REFERENCE)
    ntrain, neval = args.train, args.steps
    frames = list(inputs.trajectory(n=1000, frames=ntrain + neval, seed=1000))
    tb = 1.0 / inputs.TOANG
    d = inputs.synthetic_input(1000, 1000, 8, 590)
    modes = {}
    for name, mode, train, prec in [("reference", DDX_GUESS_REFERENCE, 0, 0),
                                    ("previous", DDX_GUESS_PREVIOUS, 0, 0),
                                    ("learned", DDX_GUESS_LEARNED, ntrain, 0),
                                    ("learned_previous", DDX_GUESS_LEARNED_PREVIOUS, ntrain, 0),
                                    ("mixed_learned_previous", DDX_GUESS_LEARNED_PREVIOUS,
                                     ntrain, 1)]:
        q, xyz, r = frames[0]
        h = DDCosmo(xyz[:, 0] * tb, xyz[:, 1] * tb, xyz[:, 2] * tb, r * tb, lmax=8, ngrid=590)
        h.set_guess(mode)
        h.set_precision(prec)
        its, t_eval, es = [], 0.0, []
        for f, (q, xyz, r) in enumerate(frames):
            h.upload(xyz[:, 0] * tb, xyz[:, 1] * tb, xyz[:, 2] * tb, r * tb, q)
            if f >= ntrain:
                t0 = time.perf_counter()
                h.run()
                t_eval += time.perf_counter() - t0
                st = h.stats()
                its.append((st.niter_direct, st.niter_adjoint))
                es.append(h.result()[0])
            else:
                h.run()
                if train:
                    h.guess_train(steps=8, lr=2e-3)
        its = np.array(its)
        modes[name] = {"solves_per_s": neval / t_eval, "ms_per_solve": t_eval / neval * 1e3,
                       "iterations_direct": float(its[:, 0].mean()),
                       "iterations_adjoint": float(its[:, 1].mean()), "energies": es}
    ref_e = np.array(modes["reference"]["energies"])
    for m in modes.values():
        m["max_rel_energy_diff_vs_reference_guess"] = float(
            np.max(np.abs(np.array(m.pop("energies")) - ref_e) / np.abs(ref_e)))
    lr_ = modes["learned"]
    out = {
        "metric": METRIC, "value": lr_["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": neval, "warmup": ntrain, "ms_per_step": lr_["ms_per_solve"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "c5: 1000 synthetic atoms, lmax 8, 590 Lebedev points, eps 78.39, "
                               "eta 0.10, conv 1e-8, energy + forces, trajectory of 0.05 A "
                               "Gaussian steps; %d training frames (warmup), %d timed frames"
                               % (ntrain, neval), "precision": args.precision, "nsph": 1000},
        "guess_modes": modes,
        "roofline": None,
    }
    if not args.no_cpu_baseline:
        q, xyz, r = frames[ntrain]
        dd = inputs.DDInput(0, 1, 8, 590, 8, 1, 78.39, 0.1, q, xyz, r)
        out["cpu_baseline"] = cpu_baseline(dd, args.cpu_cores)
    print(json.dumps(out))


def mixed_beside(h, args, sec64, es64, fx64, batch):
    """The same workload with FP32 sweeps and harmonic products inside FP64
    iterations: solves/s, its speed-up over the FP64 solves timed
    just before, iterations, and the differences of the energies and forces
    from the FP64 results."""
    h.set_precision(1)
    for _ in range(max(1, args.warmup)):
        h.run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h.run()
    sec = (time.perf_counter() - t0) / args.steps
    es, fx = h.result()
    st = h.stats()
    h.set_precision(0)
    es, es64 = np.atleast_1d(es), np.atleast_1d(es64)
    out = {"ms_per_step": sec * 1e3, "speedup_vs_fp64": sec64 / sec,
           "energy_max_rel_diff_vs_fp64": float(np.max(np.abs(es - es64) / np.abs(es64))),
           "force_max_abs_diff_vs_fp64": float(np.max(np.abs(fx - fx64))),
           "stages_ms": {"solves": st.ms_direct + st.ms_adjoint}}
    if batch:
        nd, na = h.iterations()
        out["value"] = h.count / sec
        out["iterations"] = {"direct": [int(nd.min()), int(nd.max())],
                             "adjoint": [int(na.min()), int(na.max())]}
    else:
        out["value"] = 1.0 / sec
        out["iterations"] = {"direct": st.niter_direct, "adjoint": st.niter_adjoint}
    out["unit"] = "solves/s"
    return out


def training_batches(decks, count, seed):
    """`count` batches of molecules with the sizes of `decks` drawn with
    other seeds (the same generator): data the network may learn from
    without seeing the timed molecules."""
    from ddpcm_amd import inputs
    out = []
    for b in range(count):
        out.append([inputs.synthetic_input(d.n, seed + 1000 * b + q, d.lmax, d.ngrid)
                    for q, d in enumerate(decks)])
    return out


def _upload(h, decks):
    b = [dk.bohr() for dk in decks]
    h.upload(*[np.concatenate([v[k] for v in b]) for k in range(4)],
             np.concatenate([dk.charge for dk in decks]))


def _timed_leg(h, args, iters_fn, precision):
    """Warmup, one run with every kernel class timed, then the timed steps
    with the dominant class bracketed launch by launch on the library's
    stream; returns the leg's record and its results."""
    for _ in range(args.warmup):
        h.run()
    table = kernel_table(h, iters_fn)
    dominant = max(table, key=lambda k: table[k]["ms"])
    h.kernel_timer(dominant, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h.run()
    t1 = time.perf_counter()
    launches, kms = h.kernel_timer_read(dominant)
    h.kernel_timer(dominant, False)
    es, fx = h.result()
    st = h.stats()
    nd, na = h.iterations()
    sec = (t1 - t0) / args.steps
    work = algorithmic_work(h, st, iters_fn())
    roofline = None
    if dominant in work and launches > 0:
        unit, w, per_solve = work[dominant]
        per_launch = w / per_solve
        avg_s = kms * 1e-3 / launches
        if unit == "flop":
            ach = per_launch / avg_s / 1e12
            peak = FP64_PEAK_TFLOPS if precision == 0 else FP32_PEAK_TFLOPS
            roofline = {"bound": "valu" if dominant.startswith("sweep") else "mfma",
                        "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                        "traffic": None}
            if precision:
                # gfx950 issues unpacked FP32 VALU at the FP64 rate; packed
                # FP32 is kept out of this build, so 78.6 is the ceiling of
                # the FP32 sweeps as built
                roofline["peak_unpacked_fp32"] = FP64_PEAK_TFLOPS
                roofline["frac_of_unpacked"] = ach / FP64_PEAK_TFLOPS
        else:
            ach = per_launch / avg_s / 1e9
            roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None}
        tag = "c4batch" if precision == 0 else "c4batch_mixed"
        roofline["traffic"] = traffic_from_profiles(dominant, tag)
        roofline.update({"kernel": dominant, "launches": launches,
                         "avg_launch_us": avg_s * 1e6, "work_per_launch": per_launch,
                         "work_unit": unit})
    rec = {"value": h.count / sec, "unit": "solves/s", "ms_per_step": sec * 1e3,
           "stages_ms": {"setup": st.ms_setup, "rhs": st.ms_rhs,
                         "solves": st.ms_direct + st.ms_adjoint, "forces": st.ms_forces},
           "iterations": {"direct": [int(nd.min()), int(nd.max()), float(nd.mean())],
                          "adjoint": [int(na.min()), int(na.max()), float(na.mean())]},
           "kernels": table,
           "hbm_gbs": {k: table[k].get("gbs") for k in ("jacobi", "extrap")},
           "roofline": roofline}
    return rec, es, fx


def batch_bench(args):
    """c4: the metric's comparison on the 512-molecule batch (module
    docstring)."""
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmoBatch, DDX_GUESS_LEARNED, DDX_GUESS_REFERENCE
    decks = inputs.batch_inputs()
    d = decks[0]
    h = DDCosmoBatch.from_inputs(decks)
    _upload(h, decks)

    def iters_fn():
        nd, na = h.iterations()
        return list(zip(np.diff(h.sph_off).tolist(), np.diff(h.cav_off).tolist(), nd.tolist(),
                        na.tolist()))

    # plain FP64, the reference's guess
    h.set_precision(0)
    h.set_guess(DDX_GUESS_REFERENCE)
    fp64, es64, fx64 = _timed_leg(h, args, iters_fn, 0)
    # the learned guess, trained on other molecules (untimed), then the
    # mixed precision with it on the timed molecules
    train_loss = []
    h.set_guess(DDX_GUESS_LEARNED)
    for tb in training_batches(decks, args.train_batches, 90001):
        _upload(h, tb)
        for _ in range(2):
            h.run()
            train_loss.append(h.guess_train(steps=16, lr=2e-3))
    _upload(h, decks)
    h.set_precision(1)
    mixed, es, fx = _timed_leg(h, args, iters_fn, 1)
    h.set_precision(0)
    config = {"workload": "c4: %d synthetic molecules of %d..%d atoms (%d in all), lmax %d, "
                          "%d Lebedev points, eps %.2f, eta %.2f, conv 1e-%d, energy + forces, "
                          "one batched handle" % (h.count, int(np.diff(h.sph_off).min()),
                                                  int(np.diff(h.sph_off).max()), h.nsph,
                                                  d.lmax, h.ngrid, d.eps, d.eta, d.iconv),
              "precision": "mixed (FP32 sweeps and harmonic products on the increments "
                           "and FP32 DIIS residual history; FP64 residual, Jacobi update, "
                           "iterate history, extrapolation and reductions)",
              "guess": "learned (trained on %d batches of other molecules)" % args.train_batches,
              "nsph": h.nsph, "ncav": h.ncav, "ntriples": h.stats().ntriples,
              "molecules": h.count}
    out = {
        "metric": METRIC, "value": mixed["value"], "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 sweeps, products and DIIS residuals; f64 residual, iterates",
        "data": "synthetic",
        "config": config,
        "speedup_vs_fp64": mixed["value"] / fp64["value"],
        "stages_ms": mixed["stages_ms"], "iterations": mixed["iterations"],
        "roofline": mixed["roofline"],
        "hbm_gbs": mixed["hbm_gbs"],
        "errors_vs_fp64": {
            "energy_max_rel": float(np.max(np.abs(es - es64) / np.abs(es64))),
            "force_max_abs": float(np.max(np.abs(fx - fx64)))},
        "guess_training_loss": train_loss[-1] if train_loss else None,
        "kernels": mixed["kernels"],
        "fp64": fp64,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline_batch(decks, args.cpu_cores)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--train", type=int, default=200,
                    help="c5: frames whose solutions train the guess network")
    ap.add_argument("--train-batches", type=int, default=24,
                    help="c4: batches of other molecules that train the guess network")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="precision of the timed solves; an fp64 run also times the mixed "
                         "precision beside it (\"mixed_precision\" in the JSON line)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick sweep workers per CU and the projection split by timing the "
                         "workload first (ddx_autotune)")
    ap.add_argument("--cpu-cores", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one plain 1 GiB device copy (the reference point that scales the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X; this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "c5":
        trajectory_bench(args)
        return
    if args.workload == "c4":
        batch_bench(args)
        return
    from ddpcm_amd import inputs
    from ddpcm_amd.ddx import DDCosmo, DDCosmoBatch

    batch = args.workload == "c4"
    iters_fn = None
    if batch:
        # 512 molecules of 20..800 atoms (log-uniform), lmax 6, 302 points
        decks = inputs.batch_inputs()
        d = decks[0]
        h = DDCosmoBatch.from_inputs(decks)
        b = [dk.bohr() for dk in decks]
        h.upload(*[np.concatenate([v[k] for v in b]) for k in range(4)],
                 np.concatenate([dk.charge for dk in decks]))

        def iters_fn():
            nd, na = h.iterations()
            sz = np.diff(h.sph_off)
            cz = np.diff(h.cav_off)
            return list(zip(sz.tolist(), cz.tolist(), nd.tolist(), na.tolist()))
        tag = "c4batch"
    else:
        d = inputs.config_input(args.workload)
        x, y, z, r = d.bohr()
        h = DDCosmo.from_input(d)
        h.upload(x, y, z, r, d.charge)
        tag = {"c2": "2000atoms", "c3": "20000atoms"}.get(args.workload, args.workload)
    if args.precision == "mixed":
        h.set_precision(1)
    tuned = None
    if args.autotune:
        w, sp = h.autotune()
        tuned = {"sweep_workers_per_cu": w, "projection_split": sp}
    for _ in range(args.warmup):
        h.run()
    # which kernel class dominates one solve
    table = kernel_table(h, iters_fn)
    dominant = max(table, key=lambda k: table[k]["ms"])
    # timed region; the dominant class is bracketed by HIP events on the
    # library's stream, launch by launch
    h.kernel_timer(dominant, True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h.run()
    t1 = time.perf_counter()
    launches, kms = h.kernel_timer_read(dominant)
    h.kernel_timer(dominant, False)
    es, fx = h.result()
    st = h.stats()
    sec = (t1 - t0) / args.steps
    work = algorithmic_work(h, st, iters_fn() if iters_fn else None)
    roofline = None
    if dominant in work and launches > 0:
        unit, w, per_solve = work[dominant]
        per_launch = w / per_solve
        avg_s = kms * 1e-3 / launches
        if unit == "flop":
            ach = per_launch / avg_s / 1e12
            # the FP32 operators of the mixed precision against the FP32 peak
            peak = FP64_PEAK_TFLOPS if args.precision == "fp64" else FP32_PEAK_TFLOPS
            roofline = {"bound": "mfma", "achieved": ach, "peak": peak,
                        "unit": "TFLOP/s", "frac": ach / peak, "traffic": None}
        else:
            ach = per_launch / avg_s / 1e9
            roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": ach / HBM_PEAK_GBS, "traffic": None}
        roofline["traffic"] = traffic_from_profiles(dominant, tag)
        roofline.update({"kernel": dominant, "launches": launches,
                         "avg_launch_us": avg_s * 1e6, "work_per_launch": per_launch,
                         "work_unit": unit})

    if batch:
        nd, na = h.iterations()
    mixed = None
    if args.precision == "fp64":
        mixed = mixed_beside(h, args, sec, es, fx, batch)

    parity = None
    gpath = os.path.join(ROOT, "tests", "golden", args.workload + ".npz")
    if os.path.exists(gpath) and args.workload in ("c2",):
        with np.load(gpath, allow_pickle=False) as g:
            parity = {"energy_rel_err": abs(es - float(g["esolv"])) / abs(float(g["esolv"])),
                      "force_max_abs_err": float(np.max(np.abs(fx - g["fx"]))),
                      "iters_match": bool(st.niter_direct == int(g["niter_direct"]) and
                                          st.niter_adjoint == int(g["niter_adjoint"]))}

    if batch:
        config = {"workload": "c4: %d synthetic molecules of %d..%d atoms (%d in all), lmax %d, "
                              "%d Lebedev points, eps %.2f, eta %.2f, conv 1e-%d, energy + forces, "
                              "one batched handle" % (h.count, int(np.diff(h.sph_off).min()),
                                                      int(np.diff(h.sph_off).max()), h.nsph,
                                                      d.lmax, h.ngrid, d.eps, d.eta, d.iconv),
                  "precision": args.precision, "nsph": h.nsph, "ncav": h.ncav,
                  "ntriples": st.ntriples, "molecules": h.count}
        iterations = {"direct": [int(nd.min()), int(nd.max())],
                      "adjoint": [int(na.min()), int(na.max())]}
        value = h.count / sec
    else:
        config = {"workload": "%s: %d synthetic atoms, lmax %d, %d Lebedev points, eps %.2f, "
                              "eta %.2f, conv 1e-%d, energy + forces" % (
                                  args.workload, d.n, d.lmax, h.ngrid, d.eps, d.eta, d.iconv),
                  "precision": args.precision, "nsph": d.n, "ncav": h.ncav,
                  "ntriples": st.ntriples}
        iterations = {"direct": st.niter_direct, "adjoint": st.niter_adjoint}
        value = 1.0 / sec
    out = {
        "metric": METRIC, "value": value, "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": sec * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f64" if args.precision == "fp64"
                  else "f32 sweeps, products and DIIS residuals; f64 residual, iterates"),
        "data": "synthetic",
        "config": config,
        "stages_ms": {"setup": st.ms_setup, "rhs": st.ms_rhs,
                      "solves": st.ms_direct + st.ms_adjoint, "forces": st.ms_forces},
        "iterations": iterations,
        "energy_hartree": es if not batch else float(np.sum(es)),
        "parity_vs_reference": parity,
        "roofline": roofline,
        "kernels": table,
        "hbm_gbs": {k: table[k].get("gbs") for k in ("jacobi", "extrap")},
        "mixed_precision": mixed,
        "autotune": tuned,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = (cpu_baseline_batch(decks, args.cpu_cores) if batch
                               else cpu_baseline(d, args.cpu_cores))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
