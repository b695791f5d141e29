"""Benchmark: rays/s for a 10000 x 10000 IPM magnification map on one MI355X.

Workload (the headline configuration): kappa_tot = 0.45, shear = 0.5, smooth
fraction 0.2 (kappa_star = 0.36), Kroupa masses over 0.08-1.5 M_sun, circular
star field from seed 1 (~1.28e6 stars by the reference's sizing formula),
source plane 10000 x 10000 pixels over half-length 25 theta_E, 10 rays per
pixel, FMM order p = 30, leaves of <= 32 near-field stars. A "ray" is one IPM
cell of the image-plane lattice (4.69e10 cells).

A step is one map from the resident star field: quadtree build + FMM +
shooting every cell into a zeroed map. `value` is cells per second of the
whole step, mixed precision (FP64 anchors, FP32 leaf-relative near field and
Horner, pixel-local FP32 polygons). Reported beside it, on the same stars and
lattice: the same-algorithm FP64 kernel, the reference-faithful FP64 kernel
(on a block of rows), the median relative error of the mixed map against the
FP64 map, the kernel's roofline, and the FP64 CPU oracle on sampled rows.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rays/sec, 10k² IPM map: mixed vs FP64 at ≤1e-3 median rel err"
C2 = dict(kappa_tot=0.45, shear=0.5, kappa_star=0.36, theta_star=1.0, mass_function="kroupa", m_solar=1.0,
          m_lower=0.08, m_upper=1.5, light_loss=0.001, rectangular=False, approx=True, safety_scale=1.37,
          center_y1=0.0, center_y2=0.0, half_length_y1=25.0, half_length_y2=25.0, num_pixels_y1=10000,
          num_pixels_y2=10000, num_rays_y=10, random_seed=1, write_stars=False, write_maps=False,
          write_histograms=False)
P_ORDER = 30
PEAK_VALU_TFLOPS = 78.6     # MI355X vector FP64 = unpacked FP32 (packed FP32 instructions are not used)
PEAK_HBM_GBS = 8000.0       # MI355X HBM3E
PEAK_PACKED_FP32_TFLOPS = 157.3


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_ipm_fast_10000px_traffic.json")
SQ_PROFILE = os.path.join(ROOT, "profiles", "k_ipm_fast_10000px_sq.json")
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4   # 256 CUs x 4 SIMDs at 2.4 GHz, 4 cycles per wave64 instruction


def valu_instructions():
    """VALU instructions per launch of k_ipm_fast (SQ_INSTS_VALU) from the counter
    profile of this workload, or None."""
    try:
        with open(SQ_PROFILE) as f:
            return float(json.load(f)["counters_per_launch"]["SQ_INSTS_VALU"])
    except (OSError, KeyError, ValueError):
        return None


def measured_traffic():
    """HBM bytes per launch of k_ipm_fast from the FETCH_SIZE / WRITE_SIZE
    profile of this workload, scaled by a 1 GiB copy (see that file)."""
    try:
        with open(TRAFFIC_PROFILE) as fh:
            return float(json.load(fh)["hbm_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def flop_per_cell(nbar, p):
    # near field 10 per star, far-field Horner 8 per order, 16 for the rest of the lens equation
    return 10.0 * nbar + 8.0 * p + 16.0


def oracle_mirror(m, p):
    """The FP64 oracle's tree and FMM over the same star field, lens and lattice."""
    from oracle import oracle as O
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from common import Mirror
    t0 = time.time()
    mir = Mirror(O, m, p=p)
    return mir, time.time() - t0


def cpu_baseline(mir, t_setup, seconds):
    """FP64 oracle (CPU, OpenMP) on sampled lattice rows of the same map."""
    from oracle import oracle as O
    n1, n2 = mir.num_cells
    rng = np.random.default_rng(0)
    rows_done = 0
    t_shoot = 0.0
    batch = 16
    while t_shoot < seconds:
        rows = np.sort(rng.choice(n2, size=batch, replace=False))
        t0 = time.time()
        mir.ipm_map(rows=rows)
        t_shoot += time.time() - t0
        rows_done += batch
        batch = max(16, min(512, int(batch * max(1.0, (seconds - t_shoot) / max(t_shoot, 1e-3)))))
    cells = rows_done * n1
    return dict(value=cells / t_shoot, unit="rays/s", cores=O.lib().orc_num_threads(), kind="port",
                sample=f"{rows_done} random lattice rows x {n1} cells of the same map, FP64 tree+FMM "
                       f"(tree/FMM setup {t_setup:.1f} s not timed)")


# FMM stress workload: kappa = kappa_star = 0.8, shear 0.1, no smooth matter, Salpeter 0.1-1 M_sun,
# 20000^2 pixels over half-length 50
C3 = dict(kappa_tot=0.8, shear=0.1, kappa_star=0.8, theta_star=1.0, mass_function="salpeter", m_solar=1.0,
          m_lower=0.1, m_upper=1.0, light_loss=0.001, rectangular=False, approx=True, safety_scale=1.37,
          center_y1=0.0, center_y2=0.0, half_length_y1=50.0, half_length_y2=50.0, num_pixels_y1=20000,
          num_pixels_y2=20000, num_rays_y=1, random_seed=3, write_stars=False, write_maps=False,
          write_histograms=False)
PEAK_MFMA_TFLOPS = {True: 157.3, False: 78.6}   # dense FP32 / FP64 MFMA


def m2l_stress(orders, npoints=200000):
    """M2L on FP32 and FP64 MFMA against the FP64 recurrences on the FMM stress
    workload: local-coefficient and deflection errors, M2L kernel time and MFMA
    utilisation (useful FLOP = 8 (p+1)^2 per source-target pair)."""
    from irs_amd.microlensing.IPM.ipm import IPM
    rows = []
    for p in orders:
        ref = None
        for mode in (0, 2, 3):
            m = IPM(**C3, is_double=True, precision=0, expansion_order=p, m2l_mode=mode)
            m.prepare()
            m.build()
            m.build()   # second build: warm kernels and buffers
            coeffs = m.export_coeffs("local")
            rng = np.random.default_rng(11)
            hx, hy = m.lib.get_half_length_x1(m.obj), m.lib.get_half_length_x2(m.obj)
            z = hx * rng.uniform(-1, 1, npoints) + 1j * hy * rng.uniform(-1, 1, npoints)
            w = m.deflect(z)
            if mode == 0:
                ref = (coeffs, w, z)
                log(f"[m2l] p={p}: {m.num_stars} stars, {m.tree_levels} levels")
            else:
                c0, w0, _ = ref
                scale = np.maximum(np.max(np.abs(c0), axis=1, keepdims=True), 1e-300)
                ec = (np.abs(coeffs - c0) / scale).max(axis=1)
                # deflection error relative to the far-field (local) part it perturbs: |alpha - alpha_ref| / |w - z|
                ew = np.abs(w - w0) / np.maximum(np.abs(w0 - z), 1e-300)
                f32 = m.m2l_fp32
                pairs = m.m2l_pairs
                tflops = pairs * 8.0 * (p + 1) ** 2 / m.t_m2l / 1e12
                row = {"p": p, "mfma": "fp32" if f32 else "fp64", "num_stars": m.num_stars, "pairs": pairs,
                       "m2l_ms": 1e3 * m.t_m2l, "tflops": tflops, "mfma_util": tflops / PEAK_MFMA_TFLOPS[f32],
                       "coeff_rel_err_median": float(np.median(ec)), "coeff_rel_err_max": float(ec.max()),
                       "deflection_rel_err_median": float(np.median(ew)), "deflection_rel_err_max": float(ew.max()),
                       "deflection_abs_err_max": float(np.abs(w - w0).max()),
                       "alpha_error": float(m.alpha_error)}
                log(f"[m2l] {row}")
                rows.append(row)
            m.close()
    return rows


def batch_bench(n, seed=0):
    """The heterogeneous batch: n maps drawn by draw_instances (seed), shot in one
    graph launch, against the same instances shot one launch at a time."""
    from irs_amd.microlensing.IPM.batch import IPMBatch, draw_instances
    from irs_amd.microlensing.IPM.ipm import IPM
    params = draw_instances(n, seed=seed)
    t0 = time.perf_counter()
    objs = []
    for p in params:
        m = IPM(**p)
        m.prepare()
        m.build()
        objs.append(m)
    t_build = time.perf_counter() - t0
    batch = IPMBatch()
    for m in objs:
        batch.add(m)
    log(f"[batch] {n} instances, {batch.cells:.3e} cells, prepare+build {t_build:.1f} s")
    batch.run(copy_maps=False)
    log(f"[batch] first run (autotune {batch.t_tune:.2f} s): tile_h {batch.tile_h}, blocks/CU {batch.blocks_per_cu}")
    runs = []
    for _ in range(3):
        batch.run(copy_maps=False)
        runs.append(batch.t_shoot)
    t_batch = float(np.median(runs))
    log(f"[batch] graph launch {t_batch:.3f} s ({batch.cells / t_batch:.3e} cells/s)")
    t_single = 0.0
    for m in objs:
        m.shoot()
        t_single += m.t_shoot_cells
    log(f"[batch] one launch per instance: {t_single:.3f} s of kernels ({batch.cells / t_single:.3e} cells/s)")
    out = {"instances": n, "seed": seed, "cells": batch.cells, "t_prepare_build_s": t_build,
           "t_batch_graph_s": t_batch, "batch_cells_per_s": batch.cells / t_batch,
           "t_single_kernels_s": t_single, "single_cells_per_s": batch.cells / t_single,
           "tile_h": batch.tile_h, "blocks_per_cu": batch.blocks_per_cu, "t_tune_s": batch.t_tune}
    batch.close()
    for m in objs:
        m.close()
    return out


def surrogate_bench(n_train=20000, n_cal=1024, n_test=2048, epochs=150, seed=11):
    """The learned histogram surrogate: IPM histograms of n_train + n_cal +
    n_test parameter draws (batched launches), training, entropy-threshold
    calibration on n_cal draws, and KS distances / throughput on n_test
    held-out draws against plain IPM and against the mean training histogram
    (the method without the network)."""
    from irs_amd.microlensing.Surrogate import data
    from irs_amd.microlensing.Surrogate import surrogate as S
    params = data.draw_params(n_train + n_cal + n_test, seed=seed)
    t0 = time.perf_counter()
    hist, t_ipm = data.histograms(params, log=log)
    t_data = time.perf_counter() - t0
    # the floor: the same test parameters with other star fields
    test_params = params[n_train + n_cal:]
    hist2, _ = data.histograms([dict(p, random_seed=p["random_seed"] + 7919) for p in test_params], log=log)
    x = S.features(params)
    tr, ca, te = slice(0, n_train), slice(n_train, n_train + n_cal), slice(n_train + n_cal, None)
    s = S.HistogramSurrogate(max_batch=256, seed=1)
    t0 = time.perf_counter()
    losses = s.fit(x[tr], hist[tr], epochs=epochs, log=log)
    t_train = time.perf_counter() - t0
    thr = s.calibrate(x[ca], hist[ca], ks_target=0.02, quantile=0.5)
    p = s.predict(x[te])
    ks = S.ks_distance(p, hist[te])
    acc = s.accept(p)
    mean_h = S.normalise(hist[tr]).mean(axis=0)
    ks_base = S.ks_distance(np.tile(mean_h, (len(p), 1)), hist[te])
    ks_floor = S.ks_distance(hist2, hist[te])
    t_pred = s.seconds_per_histogram(x[te])
    t_map = t_ipm / len(params)
    out = {"train": n_train, "calibration": n_cal, "test": n_test, "epochs": epochs, "t_data_s": t_data,
           "t_ipm_batched_s": t_ipm, "t_train_s": t_train, "loss_first": losses[0], "loss_last": losses[-1],
           "entropy_threshold": float(thr), "fallback_when_entropy": "above" if s.reject_above else "below",
           "accepted_fraction": float(acc.mean()),
           "ks_median": float(np.median(ks)), "ks_p95": float(np.percentile(ks, 95)), "ks_max": float(ks.max()),
           "ks_accepted_median": float(np.median(ks[acc])) if acc.any() else None,
           "ks_accepted_p95": float(np.percentile(ks[acc], 95)) if acc.any() else None,
           "ks_mean_histogram_median": float(np.median(ks_base)), "ks_mean_histogram_p95": float(np.percentile(ks_base, 95)),
           "ks_ipm_vs_ipm_median": float(np.median(ks_floor)), "ks_ipm_vs_ipm_p95": float(np.percentile(ks_floor, 95)),
           "s_per_histogram_surrogate": t_pred, "s_per_histogram_ipm_batched": t_map,
           "speedup_vs_ipm": t_map / t_pred}
    log(f"[surrogate] {out}")
    s.close()
    return out


def ccf_bench(num_stars=137, num_phi=100, seed=3, runs=3):
    """Critical curves and caustics at the reference's default configuration
    (ccf.cuh:42-62: kappa 0.3, gamma 0.3, kappa_* 0.27, equal masses, 137
    stars, 100 phases, one branch): GPU phase-sweep time (HIP events), root
    updates per second, and the FP64 oracle (direct sums, numpy) on the same
    star field on the host."""
    from irs_amd.microlensing.CCF.ccf import CCF
    from oracle import ccf as OC
    from oracle import oracle as O
    kw = dict(kappa_tot=0.3, shear=0.3, kappa_star=0.27, num_stars=num_stars, num_phi=num_phi, num_branches=1,
              random_seed=seed, write_stars=False, write_length_scales=True)
    t_ccs, t_init = [], []
    for _ in range(runs + 1):
        c = CCF(**kw)
        c.run()
        t_ccs.append(c.t_ccs)
        t_init.append(c.t_init_roots)
        res = (c.num_roots, c.max_error, c.critical_curves, c.corner)
        c.close()
    nroots, max_err, cc, corner = res
    t_ccs, t_init = float(np.median(t_ccs[1:])), float(np.median(t_init[1:]))
    sweep_updates = nroots * 2 * (num_phi // 2) * 20
    out = {"num_stars": num_stars, "num_roots": nroots, "num_phi": num_phi, "t_init_roots_s": t_init, "t_ccs_s": t_ccs,
           "root_updates_per_s": sweep_updates / t_ccs, "max_error_inv_mu": max_err}
    # the oracle on the same field (equal masses: the generation corner is the final corner)
    q = np.array([1 / abs(1 - 0.3 + 0.3), 1 / abs(1 - 0.3 - 0.3)])
    gen_corner = q / np.hypot(*q) * np.sqrt(num_stars / 0.27)
    if not os.path.exists(O.LIB_PATH):
        O.build()
    stars = O.generate_stars(num_stars, 0, tuple(gen_corner), "equal", 1.0, 1.0, 1.0, seed)
    prob = OC.Problem(stars, 0.3, 0.3, 1.0, 0.27, corner)
    t0 = time.perf_counter()
    roots, _, err = OC.solve(prob, num_phi, 1)
    t_cpu = time.perf_counter() - t0
    z = cc[..., 0] + 1j * cc[..., 1]
    worst = max(float(np.abs(z[:, k, None] - roots[None, :, k]).min(1).max()) for k in range(z.shape[1]))
    out.update({"cpu_oracle_s": t_cpu, "cpu_oracle_threads": 1, "cpu_kind": "port",
                "speedup_sweep_vs_cpu": t_cpu / (t_ccs + t_init), "max_root_distance_vs_oracle": worst,
                "oracle_max_error_inv_mu": err})
    log(f"[ccf] {out}")
    return out


def ncc_bench(num_phi=100, seed=3, runs=5):
    """Number of caustic crossings at the reference's defaults (ncc.cuh:28-35:
    1000 x 1000 pixels over +-5 Einstein radii, over_sample 2, i.e. a 4000 x
    4000 grid of sub-pixels) on the caustics of the default critical-curve run
    (137 stars). GPU time (HIP events) of the crossings and of the reduction,
    and the FP64 oracle, which adds to every pixel left of every crossing as
    the reference's kernel does, on the same caustics on the host; the maps
    must be identical."""
    import tempfile
    from irs_amd.microlensing.CCF.ccf import CCF
    from irs_amd.microlensing.NCC.ncc import NCC
    from oracle import ncc as ON
    tmp = tempfile.mkdtemp(prefix="ncc_bench_")
    c = CCF(kappa_tot=0.3, shear=0.3, kappa_star=0.27, num_stars=137, num_phi=num_phi, num_branches=1,
            random_seed=seed, write_stars=False, write_critical_curves=False, outfile_prefix=tmp + "/")
    c.run()
    c.save()
    caustics = c.caustics
    c.close()
    t_ncc, t_red = [], []
    for _ in range(runs + 1):
        n = NCC(infile_prefix=tmp + "/", write_maps=False, write_histograms=False)
        n.run()
        t_ncc.append(n.t_ncc)
        t_red.append(n.t_reduce)
        gpu_map = n.num_caustic_crossings
        n.close()
    t_ncc, t_red = float(np.median(t_ncc[1:])), float(np.median(t_red[1:]))
    stats = {}
    t0 = time.perf_counter()
    fine = ON.fine_counts(caustics, (0.0, 0.0), (5.0, 5.0), (4000, 4000), stats=stats)
    want = ON.reduce_like_reference(fine, (1000, 1000), 2)
    t_cpu = time.perf_counter() - t0
    out = {"workload": "NCC 1000x1000 px, half-length 5, over_sample 2, caustics of 137 stars x 101 phases",
           "segments": int(caustics.shape[0] * (caustics.shape[1] - 1)), "crossings": stats.get("crossings", 0),
           "reference_pixel_atomics": stats.get("pixel_updates", 0), "t_ncc_s": t_ncc, "t_reduce_s": t_red,
           "gpu_s": t_ncc + t_red, "max_count": int(gpu_map.max()), "identical_to_oracle": bool(np.array_equal(gpu_map, want)),
           "cpu_oracle_s": t_cpu, "cpu_oracle_threads": 1, "cpu_kind": "port",
           "speedup_vs_cpu": t_cpu / (t_ncc + t_red)}
    # the row pass reads the 4000 x 4000 int32 sub-pixel grid once and writes the 1000 x 1000 map
    row_pass_bytes = 4000 * 4000 * 4 + 1000 * 1000 * 4
    achieved = row_pass_bytes / t_red / 1e9
    out["roofline"] = {"kernel": "k_ncc_rows_to_map", "bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                       "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS, "bytes_per_launch": row_pass_bytes}
    log(f"[ncc] {out}")
    return out


def direct_sum_check(m, mir, npoints, seed=77):
    """Deflections at seeded lattice corners, FP64 and mixed precision, against
    the host's FP64 direct sum over every star (alpha_star.cuh:18-41 without the
    tree); the reference sizes its FMM for |error| <= alpha_error
    (ipm.cuh:532-533). Near a star the mixed path's FP32 near field adds a
    rounding term of the size of the FP32 spacing of the leaf-relative offsets
    times sum m / |z - z_i|^2; the worst point is reported with its nearest
    star and that sum."""
    from oracle import oracle as O
    rng = np.random.default_rng(seed)
    n1, n2 = m.num_ray_threads
    g = mir.grid
    I = rng.integers(0, n1 + 1, npoints)
    J = rng.integers(0, n2 + 1, npoints)
    z = (g.center_x[0] - g.half_length_x[0] + 2 * g.ray_half_sep[0] * I) + \
        1j * (g.center_x[1] - g.half_length_x[1] + 2 * g.ray_half_sep[1] * J)
    t0 = time.time()
    wd = O.deflect_direct(mir.stars, mir.lens, z)
    t_direct = time.time() - t0
    alpha = float(m.alpha_error)
    out = {"points": int(npoints), "alpha_error": alpha, "oracle_s": t_direct,
           "oracle_threads": O.lib().orc_num_threads(), "oracle": "FP64 direct sum over every star"}
    keep = m.precision
    for prec, name in ((0, "fp64"), (1, "mixed")):
        m.precision = prec
        dw = np.abs(m.deflect(z) - wd)
        k = int(np.argmax(dw))
        d2 = (mir.stars[:, 0] - z[k].real) ** 2 + (mir.stars[:, 1] - z[k].imag) ** 2
        near = np.argsort(d2)[:64]
        out[name] = {"median": float(np.median(dw)), "p99": float(np.percentile(dw, 99)), "max": float(dw.max()),
                     "above_alpha_error": int(np.sum(dw > alpha)),
                     "worst_nearest_star_distance": float(np.sqrt(d2[near[0]])),
                     "worst_sum_m_over_r2": float(np.sum(mir.stars[near, 2] / d2[near]))}
    m.precision = keep
    log(f"[bench] direct-sum check {out}")
    return out


def oracle_window(m, mir, maps, x0, y0, size=512):
    """A size x size window of the map re-shot by the FP64 oracle on the host:
    the GPU lists every lattice cell that can reach the window (mixed-precision
    corners, one pixel of margin), the oracle deflects those cells through its
    own tree and FMM and rasterises them into the window, and each GPU map is
    compared with it pixel by pixel."""
    from oracle import oracle as O
    t0 = time.time()
    rec = m.window_cells(x0, y0, size, size, margin=1.0)
    t_cells = time.time() - t0
    masks = rec[:, 2].astype(np.uint64)
    cells = int(sum(int(((masks >> np.uint64(k)) & np.uint64(1)).sum()) for k in range(63)))
    t0 = time.time()
    win = mir.ipm_window(rec, x0, y0, size, size)
    t_oracle = time.time() - t0
    out = {"window": [int(x0), int(y0), size, size], "records": int(len(rec)), "cells": cells,
           "gpu_listing_s": t_cells, "oracle_s": t_oracle, "oracle_threads": O.lib().orc_num_threads(),
           "oracle": "FP64 tree + FMM on the host, same stars and expansion order"}
    for name, mp in maps.items():
        e = np.abs(mp[y0:y0 + size, x0:x0 + size] - win) / np.maximum(np.abs(win), 1e-12)
        out[f"median_rel_err_{name}"] = float(np.median(e))
        out[f"max_rel_err_{name}"] = float(np.max(e))
    log(f"[bench] oracle window {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--ref-rows", type=int, default=4096, help="rows for the reference-faithful FP64 kernel")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--direct-points", type=int, default=20000,
                    help="lattice corners checked against the host's direct sum over every star")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-irs", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain 1 GiB device-to-device copy kernel (scale for HBM byte counters)")
    ap.add_argument("--batch", type=int, default=0,
                    help="instead of the map benchmark: this many heterogeneous maps in one graph launch")
    ap.add_argument("--surrogate", action="store_true",
                    help="instead of the map benchmark: train and evaluate the histogram surrogate")
    ap.add_argument("--ccf", action="store_true",
                    help="instead of the map benchmark: critical curves and caustics at the reference's defaults")
    ap.add_argument("--ncc", action="store_true",
                    help="instead of the map benchmark: number of caustic crossings at the reference's defaults")
    ap.add_argument("--m2l-stress", action="store_true",
                    help="instead of the map benchmark: MFMA M2L against FP64 M2L on the 20000^2 FMM stress workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the map would be sharded by "
              f"lattice rows with one sum-reduction of the maps)", file=sys.stderr)
        sys.exit(2)

    if args.m2l_stress:
        print(json.dumps({"m2l_stress": m2l_stress((10, 20, 30))}), flush=True)
        return
    if args.ccf:
        print(json.dumps({"ccf": ccf_bench()}), flush=True)
        return
    if args.ncc:
        print(json.dumps({"ncc": ncc_bench()}), flush=True)
        return
    if args.batch:
        print(json.dumps({"batch": batch_bench(args.batch)}), flush=True)
        return
    if args.surrogate:
        print(json.dumps({"surrogate": surrogate_bench()}), flush=True)
        return
    if args.copy_probe:
        # a copy kernel of known size: 1 GiB read + 1 GiB written, for scaling FETCH_SIZE / WRITE_SIZE
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        torch.cuda.synchronize()
        del x, y

    from irs_amd.microlensing.IPM.ipm import IPM

    log("[bench] C2 10000^2 IPM: generating stars")
    m = IPM(**C2, is_double=True, precision=1, expansion_order=P_ORDER)
    m.prepare()
    log(f"[bench] stars {m.num_stars}, lattice {m.num_ray_threads}")
    for w in range(args.warmup):
        m.build()
        m.shoot()
        log(f"[bench] warmup {w}: shoot {m.t_shoot_cells:.3f} s")
    step_s, shoot_s, build_s = [], [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        m.build()
        t1 = time.perf_counter()
        m.shoot()
        t2 = time.perf_counter()
        step_s.append(t2 - t0)
        build_s.append(t1 - t0)
        shoot_s.append(m.t_shoot_cells)
        log(f"[bench] step {k}: build {t1 - t0:.3f} s, shoot {m.t_shoot_cells:.3f} s (events), step {t2 - t0:.3f} s")
    cells = m.cells_shot
    nbar = m.mean_near_stars
    m.finish()
    mixed_map = m.magnifications.copy()

    out = {}
    fp64_rate = ref_rate = err_median = err_win = None
    maps = {"mixed": mixed_map}
    rng = np.random.default_rng(0)
    y0, x0 = rng.integers(0, 10000 - 512, size=2)
    if not args.no_fp64:
        m.precision = 0
        m.shoot()
        fp64_rate = cells / m.t_shoot_cells
        log(f"[bench] fp64 same-algorithm shoot {m.t_shoot_cells:.3f} s")
        m.finish()
        fp64_map = m.magnifications
        maps["fp64"] = fp64_map
        e = np.abs(mixed_map - fp64_map) / np.maximum(np.abs(fp64_map), 1e-300)
        err_median = float(np.median(e))
        err_win = float(np.median(e[y0:y0 + 512, x0:x0 + 512]))
        out["fp64_max_rel_err"] = float(np.max(e))
        log(f"[bench] mixed vs fp64 median rel err {err_median:.3e} (512^2 window {err_win:.3e}), max {np.max(e):.3e}")
        # reference-faithful FP64 kernel on seeded blocks of rows spread over the whole lattice
        n2 = m.num_ray_threads[1]
        block = 64
        starts = np.sort(np.random.default_rng(5).choice(n2 // block, size=max(1, args.ref_rows // block),
                                                        replace=False)) * block
        m.kernel = 1
        ref_cells, ref_s = 0, 0.0
        for r0 in starts:
            m.shoot(int(r0), int(r0) + block)
            ref_cells += m.cells_shot
            ref_s += m.t_shoot_cells
        ref_rate = ref_cells / ref_s
        out["reference_fp64_sample"] = (f"{len(starts)} seeded blocks of {block} lattice rows spread over all "
                                        f"{n2} rows, {ref_cells} cells")
        log(f"[bench] reference fp64 kernel: {ref_cells} cells in {ref_s:.3f} s ({len(starts)} blocks)")
        m.kernel = 0
        m.precision = 1

    ms_per_step = 1000.0 * float(np.mean(step_s))
    value = cells / float(np.mean(step_s))
    shoot_rate = cells / float(np.mean(shoot_s))
    fpc = flop_per_cell(nbar, P_ORDER)
    achieved = shoot_rate * fpc / 1e12
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32+fp64",
        "data": "synthetic",
        "config": {"workload": "C2 IPM 10000x10000 px, h_y=25, kappa=0.45, gamma=0.5, s=0.2, Kroupa 0.08-1.5, "
                               "10 rays/px, p=30, seed 1",
                   "num_stars": m.num_stars, "cells": cells, "lattice": list(m.num_ray_threads),
                   "expansion_order": P_ORDER, "tree_levels": m.tree_levels, "mean_near_stars": nbar},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_VALU_TFLOPS, "traffic": measured_traffic(),
                     "traffic_source": os.path.relpath(TRAFFIC_PROFILE, ROOT),
                     "flop_per_cell": fpc, "kernel_ms": 1000.0 * float(np.mean(shoot_s)),
                     "peak_packed_fp32": PEAK_PACKED_FP32_TFLOPS,
                     "valu_instructions": valu_instructions(),
                     "valu_issue_frac": (valu_instructions() / (VALU_ISSUE_PER_S * float(np.mean(shoot_s)))
                                         if valu_instructions() else None),
                     "valu_source": os.path.relpath(SQ_PROFILE, ROOT)},
        "shoot_rays_per_s": shoot_rate,
        "build_ms": 1000.0 * float(np.mean(build_s)),
        "fp64_rays_per_s": fp64_rate,
        "reference_fp64_rays_per_s": ref_rate,
        "speedup_vs_fp64": (shoot_rate / fp64_rate) if fp64_rate else None,
        "speedup_vs_reference_fp64": (shoot_rate / ref_rate) if ref_rate else None,
        "median_rel_err_vs_fp64": err_median,
        "median_rel_err_vs_fp64_512_window": err_win,
    }
    result.update(out)
    if not args.no_cpu:
        log("[bench] CPU oracle: tree and FMM")
        mir, t_setup = oracle_mirror(m, P_ORDER)
        log("[bench] CPU oracle: direct sums at sampled lattice corners")
        result["direct_sum_check"] = direct_sum_check(m, mir, args.direct_points)
        log("[bench] CPU oracle: 512^2 window")
        result["cpu_oracle_window"] = oracle_window(m, mir, maps, x0, y0)
        log("[bench] CPU oracle baseline")
        result["cpu_baseline"] = cpu_baseline(mir, t_setup, args.cpu_seconds)
        del mir
    else:
        result["cpu_baseline"] = None
    m.close()
    if not args.no_irs:
        # IRS on the same configuration with the IRS light loss (irs.cuh:51): rays/s of the shooting kernel
        from irs_amd.microlensing.IRS.irs import IRS
        r = IRS(**dict(C2, light_loss=0.01), expansion_order=P_ORDER)
        r.prepare()
        r.build()
        r.shoot()
        r.shoot()
        irs_rate = r.cells_shot / r.t_shoot_rays
        fpr = flop_per_cell(r.mean_near_stars, P_ORDER)
        irs_tf = irs_rate * fpr / 1e12
        result["irs"] = {"rays": r.cells_shot, "num_stars": r.num_stars, "shoot_s": r.t_shoot_rays,
                         "rays_per_s": irs_rate, "mean_near_stars": r.mean_near_stars,
                         "roofline": {"kernel": "k_irs_fast", "bound": "valu", "achieved": irs_tf,
                                      "peak": PEAK_VALU_TFLOPS, "unit": "TFLOP/s", "frac": irs_tf / PEAK_VALU_TFLOPS,
                                      "flop_per_ray": fpr,
                                      # int32 atomics that leave the LDS window, 8 B each, at most one per ray
                                      "atomic_bytes_per_ray_max": 8}}
        log(f"[bench] IRS: {r.cells_shot} rays in {r.t_shoot_rays:.3f} s")
        r.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
