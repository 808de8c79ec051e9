#!/usr/bin/env python
"""Throughput of MARTINI's source-insertion path on one MI355X.

One step = one insertion of a synthetic disc into its cube, with the particle
records already resident in HBM: zero the cube, bin and sort the particles
into cells (mi_bin), gather and reduce (mi_gather).  The metric is
particle-voxel updates per second, U = Nc * sum_p n_x(p) n_y(p) over the
reference's square mask (martini/martini.py:271-274).

The default workload is the 10^7-particle disc on a 512^2 x 256 cube, where
the BF16x3 MFMA form is the headline and the FP32 form is timed in the same
run; the 10^6-particle disc on 256^2 x 128 follows as a secondary line.  The
reference's own CPU algorithm (oracle/cpu_baseline.py) is timed beside them on
the host cores: a sample of the headline workload with all host threads, and
the 10^4-particle disc in full on one thread.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--method both|plain|mfma] [--no-cpu-baseline] [--serving N]

The serving leg (``--serving N`` requests, default 512) runs the
continuous-batching service of martini_amd.serving on a mixed request stream
and reports p50/p99 per-cube latency beside the throughput.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"plain": 157.3, "mfma": 2500.0}  # MI355X dense FP32 / BF16 (public figures)
PEAK_HBM_GBS = 8000.0  # MI355X HBM3E
PRODUCTS = {"plain": 1, "mfma": 3}  # products per logical multiply-add (bf16x3 spends 3)
DTYPE = {"plain": "fp32", "mfma": "bf16x3"}
BOUND = {"plain": "valu_fp32", "mfma": "mfma"}
GATHER_KERNEL = {"plain": "k_gather_plain", "mfma": "k_gather_mfma"}
TRAFFIC_FILE = "profiles/hbm_traffic_gather.md"
METRIC = "particle-voxel updates/s; p50/p99 cube latency; MFMA vs FP32 at 1e-5 rel err"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def updates_of(m) -> int:
    """U over the padded cube with the reference's mask semantics."""
    nx, ny, nc = m.datacube._array.shape[:3]
    x, y = m.source.pixcoords[0], m.source.pixcoords[1]
    r = np.asarray(m.sph_kernel.sm_ranges, float)
    nxp = np.clip(np.minimum(nx - 1, np.floor(x + r)) - np.maximum(0, np.ceil(x - r)) + 1, 0, None)
    nyp = np.clip(np.minimum(ny - 1, np.floor(y + r)) - np.maximum(0, np.ceil(y - r)) + 1, 0, None)
    return int(nc * np.sum(nxp * nyp))


def prepare(workload_name: str):
    """Host FP64 setup (Martini's constructor) and record packing, timed."""
    from martini_amd import Martini, _lib, _pack, synthetic

    w = synthetic.WORKLOADS[workload_name]
    src, dc, k, s = synthetic.build(w)
    t0 = time.perf_counter()
    m = Martini(source=src, datacube=dc, sph_kernel=k, spectral_model=s, quiet=True)
    t1 = time.perf_counter()
    rec, desc = _pack.pack_particles(m.source, m.sph_kernel, m.spectral_model, dc,
                                     out_scale=1.0 / dc.px_size**2)
    n_pairs = _lib.count_pairs(rec, desc)
    t2 = time.perf_counter()
    host = {"setup_s": t1 - t0, "pack_s": t2 - t1}
    return w, m, rec, desc, n_pairs, host


def upload(rec, method: str, nc=None):
    """Records into HBM as the insertion path lays them out (Inserter.upload);
    returns the device records and the host-to-device copy time in ms."""
    import torch

    from martini_amd.insertion import layout_records

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    host = torch.from_numpy(np.ascontiguousarray(rec).view(np.uint8))
    d = torch.empty(host.shape, dtype=torch.uint8, device=dev)  # allocation outside the copy
    d.copy_(host)  # the first copy from pageable memory also faults its pages in
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.copy_(host)
    torch.cuda.synchronize()
    h2d_ms = (time.perf_counter() - t0) * 1e3
    return layout_records(d, method, nc), h2d_ms


def time_gpu(d_rec, n, desc, n_pairs, method: str, steps: int, warmup: int):
    """ms per step, ms of the gather stage and of binning (HIP events on the
    launch stream), and binning status."""
    import torch

    from martini_amd import _lib

    dev = d_rec.device
    cube = torch.zeros(desc.nx * desc.ny * desc.nc, dtype=torch.float32, device=dev)
    ws = torch.empty(_lib.workspace_bytes(desc, n, n_pairs), dtype=torch.uint8, device=dev)
    stream = torch.cuda.current_stream()

    def step(ev=None):
        cube.zero_()
        if ev is not None:
            ev[0].record(stream)
        _lib.bin(d_rec, n, n_pairs, desc, ws, stream, method=method)
        if ev is not None:
            ev[1].record(stream)
        _lib.gather(d_rec, n, n_pairs, desc, cube, ws, method, stream)
        if ev is not None:
            ev[2].record(stream)

    for _ in range(warmup):
        step()
    st = _lib.status(desc, n, n_pairs, ws, stream)
    if st[1]:
        raise RuntimeError("pair capacity overflow")
    evs = [tuple(torch.cuda.Event(enable_timing=True) for _ in range(3)) for _ in range(steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for k in range(steps):
        step(evs[k])
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / steps
    bin_ms = float(np.mean([a.elapsed_time(b) for a, b, _ in evs]))
    gather_ms = float(np.mean([b.elapsed_time(c) for _, b, c in evs]))
    return ms, gather_ms, bin_ms, st, cube


def measured_traffic(workload: str, method: str):
    """HBM bytes per gather-stage launch from the committed FETCH_SIZE /
    WRITE_SIZE counter runs (profiles/hbm_traffic_gather.md), or None."""
    path = os.path.join(ROOT, TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    in_stage = False
    for line in open(path):
        if line.startswith("Gather stage"):
            in_stage = True
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if in_stage and len(cells) == 4 and cells[0] == workload and cells[1] == method:
            return int(cells[2])
    return None


def beam_leg(desc, fwhm_px: float, reps: int = 5) -> dict:
    """Time mi_convolve_beam on a cube of the workload's shape with a circular
    Gaussian beam of fwhm_px pixels (truncated at 4 FWHM, the reference's
    default): the two one-dimensional passes Martini.convolve_beam takes for
    such a separable beam, and the direct 2D sum over the whole image;
    roofline on 2 flops per tap and voxel against the FP32 peak."""
    import torch

    from martini_amd import DataCube, _lib
    from martini_amd.beams import GaussianBeam
    from martini_amd.martini import beam_factors

    b = GaussianBeam(bmaj=fwhm_px, bmin=fwhm_px, truncate=4)
    b.init_kernel(DataCube(n_px_x=1, n_px_y=1, n_channels=1, px_size=1.0, channel_width=1.0))
    K = np.ascontiguousarray(b.kernel, np.float32)
    u, v = beam_factors(b.kernel)
    px, py = b.needs_pad()
    nx, ny, nc = desc.nx + 2 * px, desc.ny + 2 * py, desc.nc
    dev = torch.device("cuda", 0)
    d_in = torch.rand((nx, ny, nc), dtype=torch.float32, device=dev)
    d_tmp, d_out = torch.empty_like(d_in), torch.empty_like(d_in)
    d_k = torch.from_numpy(K).to(dev)
    d_u = torch.from_numpy(u.astype(np.float32)).to(dev)
    d_v = torch.from_numpy(v.astype(np.float32)).to(dev)
    stream = torch.cuda.current_stream()

    def direct():
        _lib.convolve_beam(d_in, d_out, nx, ny, nc, d_k, K.shape[0], K.shape[1], stream)

    def separable():
        _lib.convolve_beam(d_in, d_tmp, nx, ny, nc, d_u, u.size, 1, stream)
        _lib.convolve_beam(d_tmp, d_out, nx, ny, nc, d_v, 1, v.size, stream)

    def timed(fn):
        fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    ms_sep, ms_dir = timed(separable), timed(direct)
    vox = float(nx) * ny * nc
    f_sep, f_dir = 2.0 * vox * (u.size + v.size), 2.0 * vox * K.shape[0] * K.shape[1]
    return {"cube": f"{nx}x{ny}x{nc} (padded)", "beam_fwhm_px": fwhm_px,
            "beam_image_px": int(K.shape[0]), "ms": ms_sep, "path": "separable (2 passes)",
            "achieved_tflops": f_sep / (ms_sep * 1e-3) / 1e12,
            "frac_fp32_peak": f_sep / (ms_sep * 1e-3) / 1e12 / PEAK_TFLOPS["plain"],
            "hbm_gbs": 4 * 4.0 * vox / (ms_sep * 1e-3) / 1e9,
            "frac_hbm_peak": 4 * 4.0 * vox / (ms_sep * 1e-3) / 1e9 / PEAK_HBM_GBS,
            "direct_ms": ms_dir,
            "direct_achieved_tflops": f_dir / (ms_dir * 1e-3) / 1e12}


def serving_leg(n_requests: int, method: str, slots: int, load: float, seed: int = 5000) -> dict:
    """The serving form on a request stream: capacity with every request
    waiting at t = 0, then the same requests with Poisson arrivals at `load`
    times that capacity.  Raw particle arrays are resident in HBM before the
    stream starts; latency runs from arrival to the cube in pinned host memory."""
    import dataclasses

    import torch

    from martini_amd import serving

    specs = serving.request_stream(n_requests, seed=seed)
    t = time.time()
    sources = [serving.disc_on_device(s) for s in specs]
    torch.cuda.synchronize()
    log(f"[bench] serving: {n_requests} requests, {sum(s.n for s in specs):.3e} particles "
        f"generated in HBM in {time.time() - t:.1f}s")
    svc = serving.Service(slots=slots, method=method, keep_cubes=False)
    svc.warm()
    at0 = [dataclasses.replace(s, arrival=0.0) for s in specs]
    svc.run(at0[:8], sources[:8])  # allocator and code warm-up
    sat = serving.summarize(svc.run(at0, sources))
    capacity = n_requests / sat["wall_s"]
    rate = load * capacity
    scale = specs[-1].arrival * rate / n_requests  # stretch the drawn gaps to the target rate
    timed = [dataclasses.replace(s, arrival=s.arrival / scale) for s in specs]
    res = serving.summarize(svc.run(timed, sources))
    log(f"[bench] serving: capacity {capacity:.1f} req/s ({sat['updates_per_s']:.3e} updates/s "
        f"saturated); at {rate:.1f} req/s p50 {res['p50_ms']:.2f} ms p99 {res['p99_ms']:.2f} ms")
    return {
        "requests": n_requests,
        "method": method,
        "slots": slots,
        "capacity_req_per_s": capacity,
        "saturated_updates_per_s": sat["updates_per_s"],
        "offered_load": load,
        "rate_req_per_s": rate,
        "p50_ms": res["p50_ms"],
        "p99_ms": res["p99_ms"],
        "mean_ms": res["mean_ms"],
        "queue_p50_ms": res["queue_p50_ms"],
        "updates_per_s": res["updates_per_s"],
        "particles": res["particles"],
        "admission_waits": svc.admission_waits,
        "mix": "N log-U[1e4,1e7], incl U(0,90), D U(3,30) Mpc, cubes {64^2x32..512^2x256}, "
               "CubicSpline/WendlandC2, thermal",
    }


def time_forms(workload: str, methods, steps: int, warmup: int) -> dict:
    """Host setup, upload and the timed steps of each method on one workload."""
    import torch

    t = time.time()
    w, m, rec, desc, n_pairs, host = prepare(workload)
    U = updates_of(m)
    n = int(rec.size)
    log(f"[bench] {workload}: {n} particles, {n_pairs} cell pairs, U={U:.3e}, "
        f"host setup {host['setup_s']:.1f}s + packing {host['pack_s']:.1f}s "
        f"({time.time() - t:.1f}s with the disc)")
    forms = {}
    for method in methods:
        d_rec, h2d_ms = upload(rec, method, desc.nc)
        ms, gather_ms, bin_ms, st, _ = time_gpu(d_rec, n, desc, n_pairs, method, steps, warmup)
        del d_rec
        torch.cuda.empty_cache()
        flops = 2.0 * U * PRODUCTS[method]
        achieved = flops / (gather_ms * 1e-3) / 1e12
        peak = PEAK_TFLOPS[method]
        traffic = measured_traffic(workload, method)
        forms[method] = {
            "ms_per_step": ms,
            "bin_ms": bin_ms,
            "gather_ms": gather_ms,
            "updates_per_s": U / (ms * 1e-3),
            "dtype": DTYPE[method],
            "work_items": int(st[2]),
            "split_items": int(st[3]),
            "h2d_ms": h2d_ms,
            "pcie_inclusive_updates_per_s": U / ((ms + h2d_ms) * 1e-3),
            "roofline": {
                "bound": BOUND[method], "kernel": GATHER_KERNEL[method] + " + k_reduce",
                "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                "traffic": traffic,
                "traffic_source": (f"{TRAFFIC_FILE} (rocprofv3 FETCH_SIZE / WRITE_SIZE runs of "
                                   "this workload and method)") if traffic is not None else None,
            },
        }
        log(f"[bench] {workload} {method}: {ms:.3f} ms/step (bin {bin_ms:.3f}, gather "
            f"{gather_ms:.3f}), {U / (ms * 1e-3):.3e} updates/s, roofline "
            f"{achieved:.1f}/{peak:.0f} TFLOP/s, items {st[2]}, split items {st[3]}, "
            f"H2D {h2d_ms:.1f} ms")
    return {"workload": w, "martini": m, "desc": desc, "U": U, "n": n, "n_pairs": n_pairs,
            "host": host, "forms": forms}


def cpu_legs(headline, budget_s: float) -> dict:
    """The reference's algorithm on the host: a sample of the headline
    workload with every host thread, and the 10^4-particle disc in full on one
    thread (its configuration's ncpu = 1)."""
    from martini_amd import synthetic
    from oracle import cpu_baseline
    from oracle import martini_oracle as O

    def prepared(w):
        a = w.arrays()
        cube = O.CubeSpec(w.n_px, w.n_px, w.n_channels, w.px_size, w.channel_width,
                          spectral_centre=w.vsys)
        K = O.adaptive(O.KernelSpec(w.kernel))
        return O.prepare(a.xyz, a.vxyz, a.mHI, a.hsm, a.T, distance_mpc=w.distance, cube=cube,
                         kernel=K, rotation=w.rotation.as_matrix())

    w = headline["workload"]
    log(f"[bench] cpu baseline: {w.name}, sampling ~{budget_s:.0f}s of the reference loop")
    main = cpu_baseline.measure(prepared(w), budget_s=budget_s)
    main.pop("seconds_extrapolated", None)
    main["workload"] = w.name
    main["host_cpus"] = os.cpu_count()
    legs = [dict(main)]
    w1 = synthetic.WORKLOADS["disc1e4_64px"]
    log(f"[bench] cpu baseline: {w1.name} in full, ncpu=1")
    c1 = cpu_baseline.run_full(prepared(w1), ncpu=1)
    c1["workload"] = w1.name
    legs.append(c1)
    main["legs"] = legs
    log(f"[bench] cpu baseline: {main['value']:.3e} updates/s ({main['cores']} threads, "
        f"{w.name}); {c1['value']:.3e} updates/s ({w1.name}, 1 thread)")
    return main


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="disc1e7_512px")
    ap.add_argument("--method", default="both", choices=["both", "plain", "mfma"],
                    help="'both': the MFMA form is the headline, the FP32 form is timed beside it")
    ap.add_argument("--secondary", default="disc1e6_256px",
                    help="workload of the secondary line, both forms ('' skips it)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="one device copy of this many MiB before timing (counter calibration)")
    ap.add_argument("--beam-fwhm-px", type=float, default=3.0,
                    help="beam FWHM (pixels) of the beam-convolution timing (0 skips it)")
    ap.add_argument("--serving", type=int, default=512,
                    help="requests in the serving leg (0 skips it)")
    ap.add_argument("--serving-method", default="mfma", choices=["plain", "mfma"])
    ap.add_argument("--serving-slots", type=int, default=4)
    ap.add_argument("--serving-load", type=float, default=0.8)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU "
              "sharding of this path is a design note, not a measured configuration.",
              file=sys.stderr)
        return 2

    if args.copy_probe_mb > 0:
        import torch

        a = torch.ones(args.copy_probe_mb << 18, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    methods = ["mfma", "plain"] if args.method == "both" else [args.method]
    head = time_forms(args.workload, methods, args.steps, args.warmup)
    lead = methods[0]
    f = head["forms"][lead]
    w, m, desc, U = head["workload"], head["martini"], head["desc"], head["U"]
    roofline = dict(f["roofline"])

    secondary = None
    if args.secondary and args.secondary != args.workload:
        sec = time_forms(args.secondary, ["mfma", "plain"], args.steps, args.warmup)
        secondary = {"workload": args.secondary, "particles": sec["n"],
                     "updates_per_step": sec["U"], "cell_pairs": sec["n_pairs"],
                     "host": sec["host"], "forms": sec["forms"]}
        del sec

    beam = beam_leg(desc, args.beam_fwhm_px) if args.beam_fwhm_px > 0 else None
    if beam is not None:
        log(f"[bench] beam convolution {beam['cube']} with a {beam['beam_image_px']} px beam "
            f"image: {beam['ms']:.3f} ms separable, {beam['direct_ms']:.3f} ms direct")

    srv = None
    if args.serving > 0:
        srv = serving_leg(args.serving, args.serving_method, args.serving_slots, args.serving_load)

    cpu = None if args.no_cpu_baseline else cpu_legs(head, args.cpu_seconds)

    forms = head["forms"]
    out = {
        "metric": METRIC,
        "value": f["updates_per_s"],
        "unit": "particle-voxel updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[lead],
        "data": "synthetic",
        "config": {
            "workload": args.workload,
            "particles": head["n"],
            "cube": f"{w.n_px}x{w.n_px}x{w.n_channels}",
            "px_arcsec": w.px_size,
            "channel_kms": w.channel_width,
            "kernel": type(m.sph_kernel).__name__,
            "spectrum": "GaussianSpectrum(thermal)",
            "method": lead,
            "updates_per_step": U,
            "cell_pairs": head["n_pairs"],
            "bin_ms": f["bin_ms"],
            "gather_ms": f["gather_ms"],
        },
        "forms": forms,
        "mfma_over_plain": (forms["plain"]["ms_per_step"] / forms["mfma"]["ms_per_step"]
                            if "plain" in forms and "mfma" in forms else None),
        "host": dict(head["host"], h2d_ms=f["h2d_ms"],
                     pcie_inclusive_updates_per_s=f["pcie_inclusive_updates_per_s"]),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
        "serving": srv,
        "beam_convolution": beam,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
