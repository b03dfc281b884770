"""miniqmc-style driver over one walker batch (src/Drivers/miniqmc.cpp).

    python -m miniqmc_amd.driver -g "1 1 2" -w 128 -n 5 -k 32 [-j] [-N 1] [-x 1.7] [-s 11] [-t fine] [--fp64]

Per step, as the reference's main loop (miniqmc.cpp:406-495): one VMC sweep of every electron of
every walker ("Diffusion": evaluateLog's refresh, the moves, completeUpdates, evaluateGL), then
the non-local pseudopotential quadrature over every electron-ion pair inside Rmax ("Pseudopotential").
At the end it prints a flat timer table and the reference's three throughput lines
(miniqmc.cpp:509-523), and writes info_<na>_<nb>_<nc>.xml with the same elements
(resources/hardware, timing, particles, run/driver; miniqmc.cpp:525-551).

Timers carry the reference's names (miniqmc.cpp:137-147, WaveFunction.cpp:38-41) and its flat-profile
line format (TimerManagerClass::print_flat, NewTimer.cpp:243-270), or with --stack-timers its stack
profile (print_stack, NewTimer.cpp:281-331: the nesting of the reference's scopes, inclusive and
exclusive time); info_*.xml holds the nested <timing> tree either way (output_timing,
NewTimer.cpp:333-395).  Host wall clock: Total, Setup,
Initialization (evaluateLog from scratch), Diffusion, Pseudopotential.  Device time of the sweep's
stages from HIP events, under the reference scope each one does the work of:
  * "New Gradient"      the SPO evaluation at every proposal (most of ratioGrad's time in the
                        reference) and the J1 terms there;
  * "Update"            the fused move kernel: evalGrad and ratioGrad's dot products, the J2 rows,
                        acceptMove and the delayed-update bookkeeping ("Current Gradient", "Set
                        active", "Make move" and "Accept move" have no separate kernel);
  * "Complete Updates"  delayed-update flush launches and the FP64 inverse refresh;
  * "Kinetic Energy"    evaluateGL.

Differences from the reference, by design:
  * every walker runs in one batch on the GPU, so there is one timer per stage instead of one per
    walker thread;
  * acceptance: the reference's rule u < r with r = 0.5 (``-r`` only accepts 0.5), or Metropolis;
  * one random sphere per walker and pseudopotential evaluation (each mover rotates its own);
  * systems: the tilings with a prepared workload, -g "1 1 1" (NiO-32) and -g "1 1 2" (NiO-64);
  * options the reference has and this driver does not take: -a (orbital tiles of the spline: a
    workgroup here evaluates all orbitals of a point), -m (mesh factor: the workloads' grids are
    fixed), -b (the reference's scalar implementations);
  * -t: every timer reported here is a coarse-level one in the reference too (miniqmc.cpp:312,
    WaveFunction.cpp:38-41), so "coarse" and "fine" print the same set.
"""
from __future__ import annotations

import argparse
import sys
import time
import xml.etree.ElementTree as ET

from miniqmc_amd import nlpp, qmcx, systems

TILINGS = {(1, 1, 1): "nio32", (1, 1, 2): "nio64"}


def parse_args(argv=None):
    ap = argparse.ArgumentParser(prog="miniqmc_amd.driver", description=__doc__.splitlines()[0])
    ap.add_argument("-g", default="1 1 1", help='3D tiling, "na nb nc"')
    ap.add_argument("-w", type=int, default=16, help="number of walkers")
    ap.add_argument("-n", type=int, default=5, help="number of MC steps")
    ap.add_argument("-k", type=int, default=32, help="delayed update rank")
    ap.add_argument("-j", action="store_true", help="enable the three-body Jastrow")
    ap.add_argument("-N", type=int, default=1, help="number of MC substeps (sweeps per step)")
    ap.add_argument("-t", default="fine", help="timer level: coarse or fine")
    ap.add_argument("-v", action="store_true", help="verbose output")
    ap.add_argument("-V", action="store_true", help="print version information and exit")
    ap.add_argument("-x", type=float, default=1.7, help="Rmax of the pseudopotential")
    ap.add_argument("-r", type=float, default=0.5, help="acceptance threshold of the reference rule (0.5 only)")
    ap.add_argument("-s", type=int, default=11, help="random seed")
    ap.add_argument("--tau", type=float, default=0.3)
    ap.add_argument("--metropolis", action="store_true", help="Metropolis acceptance instead of u < r")
    ap.add_argument("--fp64", action="store_true", help="FP64 batch (default FP32 with the FP64 refresh)")
    ap.add_argument("--system", default=None, help="a named workload instead of -g (e.g. nio8-cubic)")
    ap.add_argument("--info-dir", default=".", help="where info_<na>_<nb>_<nc>.xml is written")
    ap.add_argument("--stack-timers", action="store_true",
                    help="print the stack timer profile (the reference built with USE_STACK_TIMERS)")
    a = ap.parse_args(argv)
    try:
        a.tiling = tuple(int(t) for t in a.g.split())
    except ValueError:
        ap.error(f"-g takes three integers, got {a.g!r}")
    if len(a.tiling) != 3:
        ap.error(f"-g takes three integers, got {a.g!r}")
    if a.system is None:
        if a.tiling not in TILINGS:
            ap.error(f"tiling {a.tiling} has no prepared workload; available: {sorted(TILINGS)}")
        a.system = TILINGS[a.tiling]
    if a.r != 0.5:
        ap.error("the reference acceptance rule runs with r = 0.5 only")
    if a.w < 1 or a.n < 1 or a.k < 1 or a.N < 1:
        ap.error("-w, -n, -N and -k must be positive")
    if a.t not in ("coarse", "fine"):
        ap.error(f"Timer level should be 'coarse' or 'fine', name given: {a.t}")
    return a


# device stages of qmcx_run_sweeps -> the reference timer whose work each one does
STAGE_TIMERS = {"proposals": "New Gradient", "move": "Update", "flush": "Complete Updates",
                "recompute": "Complete Updates", "gl": "Kinetic Energy"}


def timer_table(timers: dict) -> str:
    """Flat profile in the reference's line format (TimerManagerClass::print_flat,
    Utilities/NewTimer.cpp:243-270): name, seconds, calls, seconds per call, seconds per thread
    (one host thread drives the batch)."""
    lines = ["", "Flat profile"]
    for name in sorted(timers):
        t, n = timers[name]
        lines.append("%-40s  %9.4f  %13d  %16.9f  %12.6f TIMER" % (name, t, n, t / (n + 2.220446049250313e-16), t))
    return "\n".join(lines)


# The reference's timer scopes as they nest at run time (miniqmc.cpp:330-500; WaveFunction.cpp's
# "Complete Updates" and "Kinetic Energy" run inside Diffusion): name -> parent
TIMER_PARENT = {"Setup": None, "Total": None, "Initialization": "Total", "Diffusion": "Total",
                "Current Gradient": "Diffusion", "New Gradient": "Diffusion", "Update": "Diffusion",
                "Complete Updates": "Diffusion", "Kinetic Energy": "Diffusion", "Pseudopotential": "Total",
                "Value": "Pseudopotential"}


def timer_stack(timers: dict) -> list:
    """Depth-first rows (level, name, inclusive, exclusive, calls) of the timers that ran, in the
    order the reference's scopes open (collate_stack_profile's order).  A timer outside the known
    scopes sits at the top level.  Exclusive time is the inclusive time less the children's; the
    device stages under Diffusion run on three streams at once, so their sum can exceed Diffusion's
    wall time, and the exclusive time is then 0."""
    order = list(TIMER_PARENT) + sorted(k for k in timers if k not in TIMER_PARENT)
    parent = {k: (TIMER_PARENT.get(k) if TIMER_PARENT.get(k) in timers else None) for k in timers}
    rows = []

    def visit(name, level):
        t, n = timers[name]
        kids = [k for k in order if k in timers and parent[k] == name]
        rows.append((level, name, t, max(0.0, t - sum(timers[k][0] for k in kids)), n))
        for k in kids:
            visit(k, level + 1)

    for k in order:
        if k in timers and parent[k] is None:
            visit(k, 0)
    return rows


def stack_table(timers: dict) -> str:
    """Stack profile in the reference's layout (TimerManagerClass::print_stack, NewTimer.cpp:281-331):
    names indented two spaces per level and padded to the longest, then inclusive time, exclusive
    time, calls and time per call."""
    rows = timer_stack(timers)
    width = max([len("Timer")] + [2 * lv + len(name) for lv, name, _, _, _ in rows])
    lines = ["Stack timer profile in seconds",
             "%s  %-9s  %-9s  %-10s  %-13s" % ("Timer".ljust(width), "Inclusive_time", "Exclusive_time", "Calls",
                                               "Time_per_call")]
    for lv, name, t, tx, n in rows:
        lines.append("%s  %9.4f  %9.4f  %13d  %16.9f" % ((" " * (2 * lv) + name).ljust(width), t, tx, n,
                                                       t / (n + 2.220446049250313e-16)))
    return "\n".join(lines)


def throughput(nw: int, nel: int, timers: dict) -> dict:
    """The reference's three throughput figures (miniqmc.cpp:514-522)."""
    out = {}
    for key, name, power in (("total", "Total", 3), ("diffusion", "Diffusion", 3), ("pseudopotential", "Pseudopotential", 2)):
        t = timers.get(name, (0.0, 0))[0]
        out[key] = nw * float(nel) ** power / t if t > 0 else float("nan")
    return out


def info_xml(path: str, nel: int, nsteps: int, timers: dict) -> None:
    """info_<na>_<nb>_<nc>.xml with the reference's element layout (miniqmc.cpp:525-551); <timing> is
    the timer tree of TimerManagerClass::output_timing (NewTimer.cpp:333-395): a timer's children
    under its <includes>."""
    res = ET.Element("resources")
    ET.SubElement(res, "hardware")
    timing = ET.SubElement(res, "timing")
    ET.SubElement(timing, "max_stack_level_exceeded").text = "no"
    ET.SubElement(timing, "max_timers_exceeded").text = "no"
    holders = [timing]   # holders[level]: where a timer of that level goes
    rows = timer_stack(timers)
    for i, (lv, name, t, tx, n) in enumerate(rows):
        del holders[lv + 1:]
        tm = ET.SubElement(holders[lv], "timer")
        ET.SubElement(tm, "name").text = name
        ET.SubElement(tm, "time_incl").text = f"{t:.6f}"
        ET.SubElement(tm, "time_excl").text = f"{tx:.6f}"
        ET.SubElement(tm, "calls").text = str(n)
        if i + 1 < len(rows) and rows[i + 1][0] > lv:
            holders.append(ET.SubElement(tm, "includes"))
    parts = ET.SubElement(res, "particles")
    p = ET.SubElement(parts, "particle")
    ET.SubElement(p, "name").text = "e"
    ET.SubElement(p, "size").text = str(nel)
    run = ET.SubElement(res, "run")
    drv = ET.SubElement(run, "driver")
    ET.SubElement(drv, "name").text = "miniqmc"
    ET.SubElement(drv, "steps").text = str(nsteps)
    ET.SubElement(drv, "substeps").text = "1"
    ET.ElementTree(res).write(path, xml_declaration=True, encoding="utf-8")


VERSION = "miniqmc_amd 1.0 (MI355X, gfx950)"


def main(argv=None) -> int:
    a = parse_args(argv)
    if a.V:
        print(VERSION)
        return 0
    t_setup = time.perf_counter()
    sysd = systems.make_system(a.system, seed=a.s, j3=a.j)
    prec = qmcx.FP64 if a.fp64 else qmcx.FP32
    dev = qmcx.System(sysd, precisions=(prec,))
    wb = qmcx.WalkerBatch(dev, a.w, delay_rank=a.k, precision=prec)
    wb.set_positions(systems.initial_positions(sysd, a.w, seed=a.s))
    t_setup = time.perf_counter() - t_setup
    mode = qmcx.ACCEPT_METROPOLIS if a.metropolis else qmcx.ACCEPT_REFERENCE
    pp = nlpp.NonLocalPP(sysd, rmax=a.x, seed=a.s)
    print(f"miniqmc_amd: {sysd.name}, {sysd.nel} electrons, {sysd.norb} orbitals per determinant, {a.w} walkers, "
          f"{a.n} steps of {a.N} substeps, delay rank {a.k}, {'FP64' if a.fp64 else 'FP32'}, Rmax {a.x}"
          f"{', three-body Jastrow' if a.j else ''}", flush=True)
    if a.v:
        print(f"  cell {sysd.lattice.tolist()}, {len(sysd.ion_pos)} ions, spline grid {tuple(sysd.ng)}, "
              f"delay form {'streamed' if wb.delay_form() else 'LDS-resident'}, tau {a.tau}, seed {a.s}", flush=True)

    timers = {}

    def add(name, t, n=1):
        tt, nn = timers.get(name, (0.0, 0))
        timers[name] = (tt + t, nn + n)

    add("Setup", t_setup)

    t0 = time.perf_counter()
    t = time.perf_counter()
    wb.evaluate_log()
    add("Initialization", time.perf_counter() - t)
    wb.set_kernel_timing(True)
    pairs = 0
    for step in range(a.n):
        t = time.perf_counter()
        wb.run(a.N, a.s, a.tau, mode, recompute_every=0 if a.fp64 else 1)
        add("Diffusion", time.perf_counter() - t)
        t = time.perf_counter()
        # the reference discards the ratios (NonLocalPP.hpp:115-140); they stay on the device
        _, cnt = wb.nlpp_evaluate(a.x, pp.randomize_walkers(a.w), max_pairs=8, ratios=False)
        add("Pseudopotential", time.perf_counter() - t)
        pairs += int(cnt.sum())
    add("Total", time.perf_counter() - t0)
    for stage, (ms, n) in wb.kernel_timing().items():
        if n:
            add(STAGE_TIMERS[stage], ms * 1e-3, int(n))

    print("================================== ")
    print(stack_table(timers) if a.stack_timers else timer_table(timers))
    print()
    print("========== Throughput ============ ")
    print()
    tp = throughput(a.w, sysd.nel, timers)
    print(f"Total throughput ( N_walkers * N_elec^3 / Total time ) = {tp['total']:.6e}")
    print(f"Diffusion throughput ( N_walkers * N_elec^3 / Diffusion time ) = {tp['diffusion']:.6e}")
    print(f"Pseudopotential throughput ( N_walkers * N_elec^2 / Pseudopotential time ) = {tp['pseudopotential']:.6e}")
    print(f"walker-moves/s {a.w * sysd.nel * a.n * a.N / timers['Diffusion'][0]:.4e}; "
          f"acceptance {wb.accepts().sum() / (a.w * sysd.nel * a.n * a.N):.3f}; "
          f"{pairs // a.n} electron-ion pairs per step")
    na, nb, nc = a.tiling
    path = f"{a.info_dir}/info_{na}_{nb}_{nc}.xml"
    info_xml(path, sysd.nel, a.n, timers)
    print(f"wrote {path}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
