"""Leapfrog gradient-evaluation throughput of the HMC sampler on one MI355X.

Workload (the dual-resolution 1 mm fit): highres 128x160x128 at 1 mm and
stdres 64x80x64 at 2 mm (2x2x2 block means of the same GP draw), centred
ellipsoid masks, kernel tau2=0.806 psi=0.132 nu=1.5, kriging neighbourhood
6.9 mm (W built on the device), circulant embedding 256x320x256, L=25.

A "step" is one post-burn-in HMC iteration: momentum draw, L+1 gradient
evaluations, energies, accept, Gibbs sigma draws and the thinned save.
value = sum(L+1) / wall time over the timed iterations (grad evals/s).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "leapfrog grad evals/s, 1mm dual-res, FP32 vs FP64 GPU; p50/p99 fit latency"
HBM_PEAK_GBS = 8000.0
DIMS = (128, 160, 128)
VOX = 1.0
THETA = (0.806, 0.132, 1.5)
RADIUS = 6.9
EMBED = (256, 320, 256)
L_STEPS = 25
THIN = 3       # the reference's saves every third post-burn-in iteration at 1 mm


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_problem():
    from dualres_amd import problem
    from dualres_amd.kriging import kriging_build
    t = time.perf_counter()
    p = problem.dual_problem(DIMS, VOX, THETA, seed_gp=20261202, seed_noise=(1, 2))
    log(f"[bench] synthetic images in {time.perf_counter() - t:.1f}s: n_h={p['Yh'].size} n_s={p['Ys'].size}")
    t = time.perf_counter()
    W, info = kriging_build(p["hr"], p["q_hr"], p["std"], p["q_std"], THETA, RADIUS, return_info=True)
    p["W"] = W
    p["kriging_s"] = time.perf_counter() - t
    p["kriging_unique"] = info["n_unique_solves"]
    log(f"[bench] kriging W {W.shape} nnz={W.nnz} unique solves={info['n_unique_solves']} in {p['kriging_s']:.1f}s")
    return p


def make_chain(p, precision, mode, seed=8675309):
    from dualres_amd import Chain
    c = Chain(p["ijk"], THETA, p["q_hr"], n_datasets=2, precision=precision, state_mode=mode, embed=EMBED, seed=seed)
    c.set_data(p["Yh"], p["Ys"], p["W"])
    c.add_data_to_mu()
    c.set_output_indices(p["ijk"])
    return c


def run_chain(c, p, steps, warmup, seed=8675309, captured=False):
    """warmup burn-in iterations of dual averaging, then `steps` timed
    post-burn-in iterations, through the C++ host's fit loop (dualgpmf's
    fit_model iterations: eps jitter, HMC update, Gibbs draws, saves every
    THIN-th iteration); captured replays each iteration from a HIP graph."""
    from dualres_amd.sampler import fit_native
    r = fit_native(c, p["Ys"].size, THETA[0], burnin=warmup, nsave=max(1, -(-steps // THIN)), thin=THIN, L=L_STEPS,
                   seed=seed, post_iters=steps, captured=captured)
    wall = r["wall_s"]
    return {"grad_evals_s": r["grad_evals"] / wall, "wall": wall, "grads": r["grad_evals"],
            "ms_per_iter": 1e3 * wall / r["iterations"], "p50_ms": 1e3 * r["p50_s"], "p99_ms": 1e3 * r["p99_s"],
            "mh_rate": r["mh_rate"], "eps": r["eps"]}


def precision_study(p, seeds=(1, 2, 3), burnin=200, nsave=200):
    """FP32 against FP64 posterior means at the 1 mm workload over longer
    chains: for each seed an FP32 and an FP64 fit (burnin + nsave iterations,
    every post-burn-in iteration saved), then the relative differences of
    their posterior means at the data voxels: FP32 vs FP64 on the same seed,
    and FP64 vs FP64 (FP32 vs FP32) on different seeds, the Monte-Carlo
    scale a precision difference has to be compared with."""
    from dualres_amd.sampler import fit_native
    means, stats = {}, {}
    for s in seeds:
        for prec in ("fp32", "fp64"):
            c = make_chain(p, prec, "compat", seed=1000 + s)
            r = fit_native(c, p["Ys"].size, THETA[0], burnin=burnin, nsave=nsave, thin=1, L=L_STEPS, seed=s)
            s1, _, _, ns = c.summary_read()
            c.close()
            means[(prec, s)] = s1 / max(ns, 1)
            stats[f"{prec}/{s}"] = {"mh_rate": round(r["mh_rate"], 4), "eps": round(r["eps"], 5), "saved": int(ns)}
            log(f"[bench] precision study: {prec} seed {s}: mh {r['mh_rate']:.2f}, eps {r['eps']:.4g}, {ns} saved")

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))
    same_seed = [rel(means[("fp32", s)], means[("fp64", s)]) for s in seeds]
    seed_pairs = [(a, b) for i, a in enumerate(seeds) for b in seeds[i + 1:]]
    fp64_seeds = [rel(means[("fp64", a)], means[("fp64", b)]) for a, b in seed_pairs]
    fp32_seeds = [rel(means[("fp32", a)], means[("fp32", b)]) for a, b in seed_pairs]
    m32 = sum(means[("fp32", s)] for s in seeds) / len(seeds)
    m64 = sum(means[("fp64", s)] for s in seeds) / len(seeds)
    out = {"burnin": burnin, "nsave": nsave, "seeds": list(seeds), "chains": stats,
           "fp32_vs_fp64_same_seed": [round(x, 5) for x in same_seed],
           "fp64_vs_fp64_other_seed": [round(x, 5) for x in fp64_seeds],
           "fp32_vs_fp32_other_seed": [round(x, 5) for x in fp32_seeds],
           "seed_averaged_fp32_vs_fp64": round(rel(m32, m64), 5)}
    log(f"[bench] precision study: fp32 vs fp64 same seed {out['fp32_vs_fp64_same_seed']}, fp64 other seed "
        f"{out['fp64_vs_fp64_other_seed']}, fp32 other seed {out['fp32_vs_fp32_other_seed']}, "
        f"seed-averaged fp32 vs fp64 {out['seed_averaged_fp32_vs_fp64']}")
    return out


def kernel_bytes(name, p, prec_bytes, c):
    """Algorithmic HBM bytes of one launch of the named kernel (read once + write once).

    With the stencil + boundary W layout the W kernels read the shared stencil
    once, per-row / per-voxel index words, each packed boundary entry (2-byte
    offset + weight) once, and the image-region fields they gather from once."""
    s = prec_bytes
    cs = 2 * s
    M0, M1, M2 = EMBED
    H0 = M0 // 2 + 1
    Nh = H0 * M1 * M2
    lo, hi = p["ijk"].min(0), p["ijk"].max(0)
    n0, n1, n2 = (hi + 1).tolist()
    nreg = n0 * n1 * n2
    T1 = H0 * M1 * n2
    T2 = H0 * n1 * n2
    nnz = p["W"].nnz
    n_s = p["Ys"].size
    n_h = p["Yh"].size
    if name == "z_fft_kick_drift":
        return 2 * T1 * cs + Nh * (4 * cs + s)          # T1 in/out; mu, p read; mu', p write; invlam
    if name in ("c2r_y", "r2c_y"):
        return T1 * cs + T2 * cs
    if name == "c2r_x":
        return T2 * cs + nreg * s
    ws = c.w_stats()
    if name == "r2c_x_residual":                        # mu, Yh, W^T v fields, mask -> T2
        nf = 4 if ws["stencil_layout"] else 3           # the stencil layout keeps W^T v in two fields
        return nreg * (nf * s + 1) + T2 * cs
    if ws["stencil_layout"]:
        ni, nd, nb, eb = ws["interior_rows"], ws["stencil_entries"], ws["boundary_rows"], ws["boundary_entries"]
        nt = ws["touched_voxels"]
        ntile, ncell, tvox = ws["interior_tiles"], ws["tile_cells"], ws["tile_voxels"]
        if name == "W_mu":   # interior: row words, Ys, v, stencil, R; boundary: row, rp, base, Ys, v; entries
            return (ni * (12 + 2 * s) + nd * (4 + s) + n_h * s) + (nb * (12 + 2 * s) + eb * (2 + s))
        if name == "Wt_v":   # interior: per tile its lattice box and voxels (mask in, W^T v out), the stencil;
            #                   boundary: voxel words (g, base, far rp; the quad ranges are per 16
            #                   voxels), entries, boundary v
            return (ntile * (12 + ncell * s + tvox * (1 + s)) + nd * (4 + s)) + (nt * (12 + s) + eb * (2 + s) + nb * s)
    if name == "W_mu":
        return nnz * (4 + s) + (n_s + 1) * 4 + 2 * n_s * s + n_h * s
    if name == "Wt_v":
        return nnz * (4 + s) + (n_h + 1) * 4 + n_h * 4 + n_h * s + n_s * s
    return 0


def leapfrog_micro(c, eps=0.01, reps=5):
    """The fixed-step leapfrog alone (L = 25: L + 1 gradient evaluations
    and the energies at both ends, no momentum draw, accept or Gibbs draws),
    `reps` times after one untimed run."""
    c.leapfrog(eps, L_STEPS)
    t = time.perf_counter()
    for _ in range(reps):
        c.leapfrog(eps, L_STEPS)
    dt = (time.perf_counter() - t) / reps
    return {"steps": L_STEPS, "eps": eps, "ms_per_leapfrog": round(1e3 * dt, 3),
            "grad_evals_per_s": round((L_STEPS + 1) / dt, 1)}


def profile_kernels(c, p, iters=3):
    c.profile(True)
    for _ in range(iters):
        c.update(0.01, L_STEPS, 1.0)
    ks, n = c.profile_kernels()
    c.profile(False)
    return ks, n


def measured_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (calibrated against a plain device copy), or None."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"hbm_traffic_dualres_1mm_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"].get(kernel)
    return None if k is None else int(k["traffic_bytes"])


def calibration_copy(nbytes=256 * 2 ** 20, reps=3):
    """Plain device-to-device copies of a known size through the HIP runtime;
    their FETCH_SIZE / WRITE_SIZE per dispatch calibrate the HBM counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    n = C.c_size_t(nbytes)
    if hip.hipMalloc(C.byref(a), n) or hip.hipMalloc(C.byref(b), n):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, n)
    for _ in range(reps):
        hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"[bench] calibration copy: {reps} x {nbytes} bytes device to device")


def engine_gp_draw(dims, vox, theta, seed):
    """A GP draw on the whole dims box from the engine itself: a chain's
    initial state is the circulant-embedded draw scaled by 1/10
    (MultiResParameters.h:743-759); float32, embedding 2 dims (smooth)."""
    from dualres_amd import Chain, problem
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ijk = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    emb = tuple(problem.smooth_even(2 * d) for d in dims)
    c = Chain(ijk, theta, problem.qform_axis(vox, dims), n_datasets=1, precision="fp32", state_mode="real",
              embed=emb, seed=seed)
    c.set_data(np.zeros(ijk.shape[0]))
    draw = 10.0 * c.mu()
    c.close()
    return draw.reshape(nz, ny, nx)


SERVE_SIZES = ((48, 48, 48), (64, 80, 64), (96, 112, 96), (128, 160, 128))


def serving_requests(n_req, nsave, burnin, seed=20261202, n_images=16):
    """The serving workload: highres grids cycling over 48^3, 64x80x64,
    96x112x96 and 128x160x128 at 1.5 mm, half single- and half
    dual-resolution (stdres 3 mm block means + N(0, 0.5^2)), ellipsoid masks,
    theta = (1, U(0.08, 0.2), U(0.8, 1.8)), N(0, 1) noise; FP32 fits of L = 25,
    6.9 mm neighbourhood, smooth embeddings.  Requests cycle over n_images
    images, each with its own HMC seed."""
    from dualres_amd import problem, serve
    rng = np.random.default_rng(seed)
    images = []
    t = time.perf_counter()
    for i in range(min(n_req, n_images)):
        dims = SERVE_SIZES[(i // 2) % len(SERVE_SIZES)]
        th = (1.0, float(rng.uniform(0.08, 0.2)), float(rng.uniform(0.8, 1.8)))
        f = engine_gp_draw(dims, 1.5, th, seed=100 + i)
        m = problem.ellipsoid_mask(dims, 0.45)
        hr = np.where(m, problem.nudge_zeros(f + rng.standard_normal(f.shape), m), 0.0)
        q_hr = problem.qform_axis(1.5, dims)
        ijk = problem.nonzero_ijk(hr)
        emb = tuple(problem.smooth_even(2 * int(v)) for v in (ijk.max(0) - ijk.min(0)))
        extra = {}
        if i % 2:
            sdims = tuple(d // 2 for d in dims)
            sm = problem.ellipsoid_mask(sdims, 0.45)
            s = problem.block_means(f, 2) + 0.5 * rng.standard_normal(sm.shape)
            extra = {"std": np.where(sm, problem.nudge_zeros(s, sm), 0.0),
                     "q_std": problem.qform_axis(3.0, sdims, origin=[q_hr[a, 3] + 0.75 for a in range(3)]),
                     "neighborhood": 6.9}
        images.append((hr, q_hr, th, emb, extra))
    log(f"[bench] serving workload: {len(images)} images in {time.perf_counter() - t:.1f}s")
    reqs = []
    for i in range(n_req):
        hr, qh, th, emb, extra = images[i % len(images)]
        reqs.append(serve.make_request(hr, qh, th, burnin=burnin, nsave=nsave, thin=1, L=L_STEPS, seed=1000 + i,
                                       precision="fp32", embed=emb, **extra))
    return reqs


def serving_leg(rates=(1.0, 2.0, 4.0), n_req=32, nsave=50, burnin=50, workers=0, seed=20261202):
    """Mixed fits through the fit server, admitted by a 200 GB device-byte
    budget (workers <= 0: no cap on concurrent fits besides the budget):
    Poisson arrivals at each rate; p50/p99 submit-to-result latency, fits/s,
    saved samples/s and post-burn-in grad evals/s per rate."""
    from dualres_amd import serve
    reqs = serving_requests(n_req, nsave, burnin, seed)
    rng = np.random.default_rng(seed)
    every = max(1, min(50, n_req // 10))
    out = []
    for rate in rates:
        arrivals = np.cumsum(rng.exponential(1.0 / rate, size=n_req))
        with serve.Server(200 * 10 ** 9, workers=workers) as srv:
            t0 = time.perf_counter()
            ids = []
            for k, (r, ta) in enumerate(zip(reqs, arrivals)):
                dt = ta - (time.perf_counter() - t0)
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit(r))
                if (k + 1) % every == 0:
                    log(f"[bench] serving at {rate}/s: {k + 1} of {n_req} submitted")
            res = []
            for k, i in enumerate(ids):
                res.append(srv.wait(i, timeout=1200))
                if (k + 1) % every == 0:
                    log(f"[bench] serving at {rate}/s: {k + 1} of {n_req} done")
            wall = time.perf_counter() - t0
        bad = [x["error"] for x in res if x["status"] != 0]
        if bad:
            raise RuntimeError(f"served fit failed: {bad[0]}")
        lat = np.array([x["total_s"] for x in res])
        grads = sum(x["grad_evals"] for x in res)
        out.append({"arrival_rate_per_s": rate, "requests": n_req, "p50_s": round(float(np.percentile(lat, 50)), 4),
                    "p99_s": round(float(np.percentile(lat, 99)), 4), "fits_per_s": round(n_req / wall, 3),
                    "saved_samples_per_s": round(n_req * nsave / wall, 1),
                    "post_burnin_grad_evals_per_s": round(grads / wall, 1)})
        log(f"[bench] serving at {rate}/s: p50 {out[-1]['p50_s']} s, p99 {out[-1]['p99_s']} s, "
            f"{out[-1]['fits_per_s']} fits/s")
    return {"workers": "budget" if workers <= 0 else workers, "budget_gb": 200,
            "workload": "highres 48^3 / 64x80x64 / 96x112x96 / 128x160x128 @1.5mm, 50% dual (stdres 3mm), "
                        f"theta (1, U(0.08,0.2), U(0.8,1.8)), fp32, {burnin} burn-in + {nsave} saved iterations, "
                        "L=25", "rates": out}


def batch_problems(n_chains=64, seed=7):
    """The batched-chains workload: n_chains dual-resolution images, highres
    dims drawn per axis from U{32..112} at 1.5 mm (stdres 3 mm block means),
    their own ellipsoid masks, theta = (1, U(0.08, 0.2), U(0.8, 1.8)) in the
    (tau^2, psi, nu) order, kriging W built on the device with a 6.9 mm
    neighbourhood, embedding per axis the smallest even 2,3,5,7-smooth length
    >= 2 max ijk."""
    from dualres_amd import problem
    from dualres_amd.kriging import kriging_build
    rng = np.random.default_rng(seed)
    out = []
    t = time.perf_counter()
    for i in range(n_chains):
        dims = tuple(int(rng.integers(32, 113)) for _ in range(3))
        th = (1.0, float(rng.uniform(0.08, 0.2)), float(rng.uniform(0.8, 1.8)))
        q = problem.dual_problem(dims, 1.5, th, seed_gp=500 + i)
        q["W"] = kriging_build(q["hr"], q["q_hr"], q["std"], q["q_std"], th, 6.9)
        q["theta"] = th
        q["embed"] = tuple(problem.smooth_even(2 * int(m)) for m in q["ijk"].max(0))
        out.append(q)
    log(f"[bench] batch workload: {n_chains} images and kriging builds in {time.perf_counter() - t:.1f}s")
    return out


def tuner_study(probs, iters=6):
    """Per chain, the device time of `iters` HMC iterations (L = 25) with the
    built-in FFT pass shapes and with the autotuner's, on its own stream."""
    from dualres_amd import Chain
    out = []
    for q in probs:
        c = Chain(q["ijk"], q["theta"], q["q_hr"], n_datasets=2, precision="fp32", state_mode="compat",
                  embed=q["embed"], seed=11)
        c.set_data(q["Yh"], q["Ys"], q["W"])
        c.add_data_to_mu()
        c.set_sigma_sq_inv([1.0, 1.0])
        ts = []
        for tuned in (False, True):
            if tuned:
                prm, _ = c.autotune()
            c.update(0.04, L_STEPS, 0.0)
            t = time.perf_counter()
            for _ in range(iters):
                c.update(0.04, L_STEPS, 0.0)
            ts.append((time.perf_counter() - t) / iters)
        out.append({"embed": list(q["embed"]), "default_ms": round(1e3 * ts[0], 3), "tuned_ms": round(1e3 * ts[1], 3),
                    "params": prm})
        c.close()
    sp = [o["default_ms"] / o["tuned_ms"] for o in out]
    log(f"[bench] tuner: single-chain speed-up median {np.median(sp):.3f} (min {min(sp):.3f}, max {max(sp):.3f})")
    return {"chains": out, "speedup_median": round(float(np.median(sp)), 3)}


def batch_leg(probs, modes=("graph", "streams", "eager"), burnin=30, timed=30, seed=900, tune=False):
    """Post-burn-in grad evals/s of all chains fitted in lockstep through one
    Batch (fit_batch: per chain dual averaging, Gibbs draws and saves; FP32
    gradients, FP64 energies), L = 25, in each batch mode; a chain's posterior
    mean must be the same bit for bit in every mode."""
    from dualres_amd import Chain
    from dualres_amd.batch import fit_batch
    from dualres_amd.sampler import HMCParameters
    out = {"chains": len(probs), "workload": "64 dual-res images, highres U{32..112}^3 @1.5mm (stdres 3mm), "
           "theta (1, U(0.08,0.2), U(0.8,1.8)), 6.9mm neighbourhood, smooth embeddings, fp32 gradients, fp64 "
           f"energies, L=25, {burnin} burn-in + {timed} timed iterations", "autotuned": tune, "modes": {}}
    means = {}
    import tempfile
    cache = os.path.join(tempfile.mkdtemp(prefix="dualres_tune_"), "fft_tuning.json")
    for mode in modes:
        chains = []
        t = time.perf_counter()
        for q in probs:
            c = Chain(q["ijk"], q["theta"], q["q_hr"], n_datasets=2, precision="fp32", state_mode="compat",
                      embed=q["embed"], seed=11)
            c.set_data(q["Yh"], q["Ys"], q["W"])
            c.add_data_to_mu()
            c.set_output_indices(q["ijk"])
            if tune:
                c.set_sigma_sq_inv([1.0, 1.0])
                c.autotune(cache)   # timed once per configuration, then read from the cache
            chains.append(c)
        if tune:
            log(f"[bench] batch {mode}: chains set up and tuned in {time.perf_counter() - t:.1f}s")
        hmcs = [HMCParameters(burnin=burnin, n_save=timed, thin=1, integrator_steps=L_STEPS,
                              starting_eps=1.0 / L_STEPS) for _ in probs]

        def tick(it, timed_it):
            if it % 10 == 0:
                log(f"[bench] batch {mode}: iteration {it}")
        summ, stats, info = fit_batch(chains, hmcs, [1.0] * len(probs), [q["Ys"].size for q in probs],
                                      [seed + i for i in range(len(probs))], mode=mode, timer=tick)
        rate = info["post_burnin_grad_evals"] / info["wall_s"]
        out["modes"][mode] = {"post_burnin_grad_evals_per_s": round(rate, 1), "wall_s": round(info["wall_s"], 3),
                              "ms_per_iteration": round(1e3 * info["wall_s"] / timed, 3)}
        means[mode] = [s.mode_mu() for s in summ]
        log(f"[bench] batch {mode}: {rate:.1f} post-burn-in grad evals/s, {1e3 * info['wall_s'] / timed:.2f} ms "
            f"per iteration of {len(probs)} chains")
        for c in chains:
            c.close()
    ref = means[modes[0]]
    out["bitwise_equal_across_modes"] = all(all(np.array_equal(a, b) for a, b in zip(ref, means[m])) for m in modes)
    best = max(out["modes"], key=lambda m: out["modes"][m]["post_burnin_grad_evals_per_s"])
    out["best_mode"] = best
    out["post_burnin_grad_evals_per_s"] = out["modes"][best]["post_burnin_grad_evals_per_s"]
    out["alone"] = alone_study(probs)
    out["alone"]["overlap_factor"] = round(out["alone"]["sum_ms"] / out["modes"][best]["ms_per_iteration"], 3)
    log(f"[bench] batch: the chains' iterations one at a time sum to {out['alone']['sum_ms']:.1f} ms, "
        f"{out['alone']['overlap_factor']:.2f}x the batch's")
    out["tuner"] = tuner_study(probs[:8])
    return out


def alone_study(probs, iters=3):
    """Each chain's HMC iteration (L = 25, eager launches) timed alone on the
    device, one chain at a time; the sum is what a batch would take with no
    overlap between chains."""
    from dualres_amd import Chain
    ms = []
    for q in probs:
        c = Chain(q["ijk"], q["theta"], q["q_hr"], n_datasets=2, precision="fp32", state_mode="compat",
                  embed=q["embed"], seed=11)
        c.set_data(q["Yh"], q["Ys"], q["W"])
        c.add_data_to_mu()
        c.set_sigma_sq_inv([1.0, 1.0])
        c.update(0.04, L_STEPS, 0.0)
        t = time.perf_counter()
        for _ in range(iters):
            c.update(0.04, L_STEPS, 0.0)
        ms.append(1e3 * (time.perf_counter() - t) / iters)
        c.close()
    return {"sum_ms": round(float(sum(ms)), 2), "median_ms": round(float(np.median(ms)), 3),
            "max_ms": round(float(max(ms)), 3), "min_ms": round(float(min(ms)), 3)}


def mce_leg(n_images=32, seed=5, nbins=64, maxdist=20.0):
    """RBF parameter estimation on a batch of synthetic 128x160x128 images at
    1 mm: each a GP draw with (tau^2, psi, nu) ~ U(0.5,1.5) x U(0.05,0.3) x
    U(0.5,1.9) on the ellipsoid mask plus N(0,1) noise.  The draws come from
    the engine itself: a chain's initial state is a circulant-embedded GP
    draw scaled by 1/10 (MultiResParameters.h:743-759).  Per image, timed:
    the binned empirical covariance (nbins bins to maxdist mm, FP64 bin sums)
    and its constrained minimum-contrast fit (psi <= nu), and the per-offset
    path estimate_rbf takes (25,348 offsets, then the same fit).  Parameter
    errors: binned against per-offset estimates and both against the truth."""
    from dualres_amd import Chain, problem
    from dualres_amd.mce import mce_binned, mce_summary, rbf_fit
    dims, vox = (128, 160, 128), 1.0
    mask = problem.ellipsoid_mask(dims, 0.45)
    ijk = problem.nonzero_ijk(mask.astype(np.float64))
    lo = ijk.min(0)
    q = problem.qform_axis(vox, dims)
    flat = np.flatnonzero(mask.ravel())
    rng = np.random.default_rng(seed)
    tb, tp, thetas = [], [], []
    t0 = time.perf_counter()
    for k in range(n_images):
        th = (float(rng.uniform(0.5, 1.5)), float(rng.uniform(0.05, 0.3)), float(rng.uniform(0.5, 1.9)))
        c = Chain(ijk - lo, th, q, n_datasets=1, precision="fp32", state_mode="real", embed=EMBED, seed=7000 + k)
        c.set_data(np.zeros(ijk.shape[0]))
        draw = 10.0 * c.mu()
        c.close()
        img = np.zeros(mask.size)
        img[flat] = draw + rng.standard_normal(flat.size)
        img = img.reshape(mask.shape)
        t = time.perf_counter()
        bc, bd, bn = mce_binned(img, q, nbins=nbins, maxdist=maxdist)
        theta_b, rc_b = rbf_fit(bc, bd, bn, constrained=True)
        tb.append(time.perf_counter() - t)
        t = time.perf_counter()
        cov, dist, npairs = mce_summary(img, q)
        theta_p, rc_p = rbf_fit(cov, dist, npairs, constrained=True)
        tp.append(time.perf_counter() - t)
        thetas.append((th, theta_b.tolist(), theta_p.tolist(), rc_b, rc_p))
        if (k + 1) % 8 == 0:
            log(f"[bench] estimation: {k + 1} of {n_images} images")
    wall = time.perf_counter() - t0
    T = np.array([t[0] for t in thetas])
    B = np.array([t[1] for t in thetas])
    P = np.array([t[2] for t in thetas])

    def med_rel(a, b):
        return [round(float(np.median(np.abs(a[:, j] - b[:, j]) / np.abs(b[:, j]))), 4) for j in range(3)]
    out = {"images": n_images, "workload": f"{dims[0]}x{dims[1]}x{dims[2]} @1mm ellipsoid-masked GP draws + N(0,1), "
           f"theta ~ U(0.5,1.5) x U(0.05,0.3) x U(0.5,1.9); {nbins} bins to {maxdist:g} mm; constrained fit",
           "binned": {"images_per_s": round(n_images / sum(tb), 3), "p50_ms": round(1e3 * float(np.median(tb)), 2),
                      "p99_ms": round(1e3 * float(np.percentile(tb, 99)), 2)},
           "per_offset": {"images_per_s": round(n_images / sum(tp), 3),
                          "p50_ms": round(1e3 * float(np.median(tp)), 2)},
           "median_rel_error": {"binned_vs_per_offset": med_rel(B, P), "binned_vs_truth": med_rel(B, T),
                                "per_offset_vs_truth": med_rel(P, T)},
           "fit_status_ok": int(sum(1 for t in thetas if t[3] == 0 and t[4] == 0)),
           "wall_s_including_draws": round(wall, 1)}
    log(f"[bench] estimation: binned {out['binned']['images_per_s']} images/s, per-offset "
        f"{out['per_offset']['images_per_s']} images/s; median rel. error binned vs per-offset "
        f"{out['median_rel_error']['binned_vs_per_offset']}")
    return out


def host_cores():
    """What the host offers this job: logical CPUs, the CPUs this process may
    run on, physical cores (distinct physical id / core id pairs), and the
    thread count the CPU baseline uses (the CPUs this process may run on,
    at most 16: one GPU's share of the host)."""
    n_logical = os.cpu_count() or 1
    try:
        n_affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        n_affinity = n_logical
    phys = set()
    try:
        pid = cid = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("physical id"):
                    pid = line.split(":")[1].strip()
                elif line.startswith("core id"):
                    cid = line.split(":")[1].strip()
                elif not line.strip() and pid is not None:
                    phys.add((pid, cid))
                    pid = cid = None
    except OSError:
        pass
    threads = max(1, min(16, n_affinity))
    return {"nproc": n_logical, "affinity": n_affinity, "physical_cores": len(phys) or None, "threads": threads,
            "threads_80pct": max(1, threads * 4 // 5)}


CFG1_THETA = (0.806, 0.132, 1.0)
CFG1_EMBED = (80, 96, 80)


def config1_problem():
    """The plain single-resolution workload: 40x48x40 at 2 mm, ellipsoid
    mask, one GP draw (tau2 = 0.806, psi = 0.132, nu = 1) + N(0, 1) noise."""
    from dualres_amd import problem
    return problem.single_problem((40, 48, 40), 2.0, CFG1_THETA, seed_gp=20261202, seed_noise=1)


def config1_leg(seed=8675309):
    """A whole single-resolution fit at the 2 mm workload, as dualgpm runs it:
    FP64, embedding 80x96x80, L = 25, 200 burn-in + 200 saved iterations
    (thin 1) through the C++ fit loop; wall time of the whole fit and the
    post-burn-in grad evals/s."""
    from dualres_amd import Chain
    from dualres_amd.sampler import fit_native
    p = config1_problem()
    c = Chain(p["ijk"], CFG1_THETA, p["q_hr"], n_datasets=1, precision="fp64", state_mode="compat",
              embed=CFG1_EMBED, seed=seed)
    c.set_data(p["Yh"])
    c.add_data_to_mu()
    c.set_output_indices(p["ijk"])
    t = time.perf_counter()
    r = fit_native(c, 0, CFG1_THETA[0], burnin=200, nsave=200, thin=1, L=L_STEPS, seed=seed)
    wall = time.perf_counter() - t
    c.close()
    out = {"workload": "single-res 40x48x40 @2mm, embedding 80x96x80, fp64, L=25, 200 burn-in + 200 saved",
           "fit_wall_s": round(wall, 3), "post_burnin_grad_evals_per_s": round(r["grad_evals"] / r["wall_s"], 1),
           "post_burnin_ms_per_iteration": round(1e3 * r["wall_s"] / r["iterations"], 3), "mh_rate": r["mh_rate"],
           "eps": r["eps"]}
    log(f"[bench] config 1: whole fit {wall:.2f}s, {out['post_burnin_grad_evals_per_s']:.0f} post-burn-in "
        f"grad evals/s, {out['post_burnin_ms_per_iteration']:.2f} ms per iteration")
    return out


def cpu_config1_chain(eps, threads, iters=400):
    """The C++ CPU restatement running the 2 mm workload's 400 HMC iterations
    (L = 25, FP64) at a fixed step size on `threads` threads."""
    from oracle import cpu_port
    p = config1_problem()
    c = cpu_port.CpuChain(p["ijk"], CFG1_THETA, p["q_hr"], CFG1_EMBED, p["Yh"], precision="fp64", threads=threads,
                          seed=42)
    c.add_data_to_mu()
    c.set_sigma_sq_inv([1.0])
    t = time.perf_counter()
    for i in range(iters):
        c.update(eps, L_STEPS)
        if (i + 1) % 50 == 0:
            log(f"[bench] cpu config 1: {i + 1} of {iters} iterations, {time.perf_counter() - t:.1f}s")
    wall = time.perf_counter() - t
    c.close()
    return {"iterations": iters, "threads": threads, "wall_s": round(wall, 2),
            "grad_evals_s": round(iters * (L_STEPS + 1) / wall, 2)}


def cpu_iterations(p, precision, threads, w_threads, iters, eps, tag):
    """Seconds of each of `iters` HMC iterations (L = 25, momentum, energies,
    accept, Gibbs draws) of the C++ restatement of the reference's sampler
    (oracle/cpu_hmc.cpp) at the benchmark workload."""
    from oracle import cpu_port
    t = time.perf_counter()
    c = cpu_port.CpuChain(p["ijk"], THETA, p["q_hr"], EMBED, p["Yh"], p["Ys"], p["W"], precision=precision,
                          threads=threads, w_threads=w_threads, seed=42)
    c.add_data_to_mu()
    c.set_sigma_sq_inv([1.0, 4.0])
    log(f"[bench] cpu {tag}: set-up {time.perf_counter() - t:.1f}s")
    out = []
    for i in range(iters):
        t = time.perf_counter()
        c.update(eps, L_STEPS)
        out.append(time.perf_counter() - t)
        log(f"[bench] cpu {tag}: iteration {i + 1}/{iters} {out[-1]:.1f}s")
    c.close()
    return out


CPU_SAMPLE = ("post-burn-in HMC iterations (L = 25: momentum draw, 26 gradient evaluations, energies, accept, "
              "Gibbs draws) of the C++17/OpenMP restatement of the reference's sampler (complex state, 4 complex 3D "
              "DFTs per leapfrog step with its own radix-4/2/3/5/7 DFT, W and W^T products on one thread) at the "
              "same workload")


def cpu_baseline(p, eps):
    """One post-burn-in iteration of the C++ CPU restatement in float32 on this
    job's share of the host cores, plus the committed study (float32 and
    float64, all and 80% of the threads, OpenMP W, medians of 3 iterations)."""
    hc = host_cores()
    t = cpu_iterations(p, "fp32", hc["threads"], 1, 1, eps, "fp32")[0]
    out = {"value": (L_STEPS + 1) / t, "unit": "grad evals/s", "cores": hc["threads"], "kind": "port",
           "sample": f"1 of the {CPU_SAMPLE}, float32, {hc['threads']} threads, {t:.1f}s",
           "host": hc}
    path = os.path.join(ROOT, "profiles", "cpu_baseline_dualres_1mm.json")
    if os.path.exists(path):
        with open(path) as f:
            out["study"] = json.load(f)
    return out


def cpu_leapfrog(p, precision, threads, eps=0.01):
    """One fixed-step L = 25 leapfrog of the C++ CPU restatement (no
    momentum draw, accept or Gibbs draws) at the benchmark workload."""
    from oracle import cpu_port
    c = cpu_port.CpuChain(p["ijk"], THETA, p["q_hr"], EMBED, p["Yh"], p["Ys"], p["W"], precision=precision,
                          threads=threads, w_threads=1, seed=42)
    c.add_data_to_mu()
    c.set_sigma_sq_inv([1.0, 4.0])
    t = time.perf_counter()
    c.leapfrog(eps, L_STEPS)
    dt = time.perf_counter() - t
    c.close()
    log(f"[bench] cpu {precision} leapfrog alone: {dt:.1f}s")
    return {"precision": precision, "threads": threads, "seconds": round(dt, 2),
            "grad_evals_s": round((L_STEPS + 1) / dt, 3)}


def cpu_study(p, eps, iters=3):
    """Medians of `iters` iterations: float32 and float64, all threads and 80%
    of them (the reference's default share, defines.h:73-74) with W on one
    thread, and all threads with OpenMP W; then a whole 400-iteration chain
    at the 2 mm single-resolution workload."""
    hc = host_cores()
    runs = []
    for prec in ("fp32", "fp64"):
        for threads, w in ((hc["threads"], 1), (hc["threads_80pct"], 1), (hc["threads"], hc["threads"])):
            ts = cpu_iterations(p, prec, threads, w, iters, eps, f"{prec} {threads} threads, W on {w}")
            med = float(np.median(ts))
            runs.append({"precision": prec, "threads": threads, "w_threads": w, "iteration_s": [round(x, 2) for x in ts],
                         "median_s": round(med, 2), "grad_evals_s": round((L_STEPS + 1) / med, 3)})
    lfs = [cpu_leapfrog(p, prec, hc["threads"]) for prec in ("fp32", "fp64")]
    cfg1 = cpu_config1_chain(eps=0.1, threads=hc["threads"])
    return {"host": hc, "sample": f"median of {iters} {CPU_SAMPLE}", "eps": eps, "runs": runs,
            "leapfrog_microbench": lfs,
            "config1_chain": dict(cfg1, sample="400 HMC iterations (L = 25, FP64, fixed step size 0.1) of the same "
                                               "restatement at the 2 mm single-resolution workload")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--no-serve", action="store_true", help="skip the fit-server latency leg")
    ap.add_argument("--serve-requests", type=int, default=32, help="requests per arrival rate in the serving leg")
    ap.add_argument("--serve-rates", default="1,2,4", help="Poisson arrival rates (fits/s) of the serving leg")
    ap.add_argument("--serve-nsave", type=int, default=50, help="saved (and burn-in) iterations per served fit")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device-to-device copy (reference for HBM counters)")
    ap.add_argument("--main-only", action="store_true",
                    help="run only the --precision complex-state chain (for profiler runs)")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched-chains leg and print its JSON")
    ap.add_argument("--mce-only", action="store_true", help="run only the parameter-estimation leg and print its JSON")
    ap.add_argument("--mce-images", type=int, default=32, help="images in the parameter-estimation leg")
    ap.add_argument("--batch-modes", default="graph,streams,eager", help="batch modes to time, in order")
    ap.add_argument("--tune", action="store_true", help="autotune the batched chains' FFT pass shapes before fitting")
    ap.add_argument("--cpu-study", action="store_true",
                    help="time only the CPU restatement (float32 and float64, all and 80%% of the threads, OpenMP W; "
                         "medians of 3 iterations) and print its JSON")
    ap.add_argument("--precision-study", action="store_true",
                    help="FP32 vs FP64 posterior means over 200 + 200 iterations on three seeds; print its JSON")
    ap.add_argument("--config1-only", action="store_true",
                    help="run only the 2 mm single-resolution fit on the GPU and the CPU restatement's 400-iteration "
                         "chain at the same workload, and print their JSON")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving leg and print its JSON (latency study with --serve-requests)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); chains and requests shard across GPUs "
              f"without communication, so the per-GPU figure is the whole measurement.", file=sys.stderr)
        sys.exit(2)
    if a.serve_only:
        print(json.dumps(serving_leg(rates=tuple(float(x) for x in a.serve_rates.split(",")),
                                     n_req=a.serve_requests, nsave=a.serve_nsave, burnin=a.serve_nsave)), flush=True)
        return
    if a.config1_only:
        g = config1_leg()
        cpu = cpu_config1_chain(eps=0.1, threads=host_cores()["threads"])
        print(json.dumps({"gpu": g, "cpu": cpu, "host": host_cores()}), flush=True)
        return
    p = build_problem()
    if a.cpu_study:
        print(json.dumps(cpu_study(p, eps=0.54)), flush=True)
        return
    if a.precision_study:
        print(json.dumps(precision_study(p)), flush=True)
        return
    if a.mce_only:
        print(json.dumps(mce_leg(a.mce_images)), flush=True)
        return
    if a.batch_only:
        print(json.dumps(batch_leg(batch_problems(), modes=tuple(a.batch_modes.split(",")),
                                   tune=a.tune)), flush=True)
        return
    results = {}
    main_chain = None
    means, mc_var = {}, {}
    for prec in ((a.precision,) if a.main_only else ("fp32", "fp64")):
        c = make_chain(p, prec, "compat")
        r = run_chain(c, p, a.steps, a.warmup)
        log(f"[bench] {prec} compat: {r['grad_evals_s']:.1f} grad evals/s, {r['ms_per_iter']:.2f} ms/iter, "
            f"mh {r['mh_rate']:.2f}, eps {r['eps']:.3g}")
        results[prec] = r
        s1, s2, _, ns = c.summary_read()
        means[prec] = s1 / max(ns, 1)
        # variance of the saved-sample mean, taken as independent draws
        mc_var[prec] = np.maximum(s2 / max(ns, 1) - means[prec] ** 2, 0.0) / max(ns, 1)
        if prec == a.precision:
            main_sum = s1.copy()
            main_chain = c
            lf = leapfrog_micro(c)
            log(f"[bench] {prec} leapfrog alone: {lf['grad_evals_per_s']:.1f} grad evals/s")
            ks, nk = profile_kernels(c, p)
        else:
            c.close()
    mean_rel = mc_ratio = None
    if len(means) == 2:
        d = means["fp32"] - means["fp64"]
        mean_rel = float(np.linalg.norm(d) / np.linalg.norm(means["fp64"]))
        # the two chains adapt their step sizes apart and so take different
        # accept decisions: their saved means differ by Monte-Carlo error.
        # ratio ~ 1: the difference is what two independent chains' would be
        mc_ratio = float(np.sqrt(np.sum(d * d) / np.sum(mc_var["fp32"] + mc_var["fp64"])))
        log(f"[bench] fp32 vs fp64 posterior mean over the saved samples: relative difference {mean_rel:.2e}, "
            f"{mc_ratio:.2f} x the Monte-Carlo spread of two independent chains")
    # the same fit with every iteration replayed from a captured HIP graph
    cc = make_chain(p, a.precision, "compat")
    rc = run_chain(cc, p, a.steps, a.warmup, captured=True)
    rc["bitwise_equal_to_eager"] = bool(np.array_equal(cc.summary_read()[0], main_sum))
    cc.close()
    log(f"[bench] {a.precision} compat, captured iterations: {rc['grad_evals_s']:.1f} grad evals/s, "
        f"{rc['ms_per_iter']:.2f} ms/iter, same run bit for bit: {rc['bitwise_equal_to_eager']}")
    rr = None
    if not a.main_only:
        cr = make_chain(p, a.precision, "real")
        rr = run_chain(cr, p, a.steps, a.warmup)
        cr.close()
        log(f"[bench] {a.precision} real: {rr['grad_evals_s']:.1f} grad evals/s")
    # roofline of the dominant kernel
    sb = 4 if a.precision == "fp32" else 8
    dom = max(ks, key=ks.get)
    byt = kernel_bytes(dom, p, sb, main_chain)
    ach = byt / (ks[dom] * 1e-3) / 1e9
    step_ms = sum(ks.values())
    log("[bench] kernels (ms/step): " + ", ".join(f"{k}={v:.4f}" for k, v in ks.items()))
    fft_names = [k for k in ks if k.startswith(("c2r", "r2c", "z_fft"))]
    fft_bytes = sum(kernel_bytes(k, p, sb, main_chain) for k in fft_names)
    fft_ms = sum(ks[k] for k in fft_names)
    wst = main_chain.w_stats()
    if not wst["stencil_layout"]:
        log("[bench] warning: W runs on the plain CSR kernels (the stencil layout does not fit this problem); "
            "the byte model above is the CSR one")
    main_chain.close()
    if a.calib_copy:
        calibration_copy()
    serving = None if a.main_only or a.no_serve else serving_leg(
        rates=tuple(float(x) for x in a.serve_rates.split(",")), n_req=a.serve_requests, nsave=a.serve_nsave,
        burnin=a.serve_nsave)
    batch = None if a.main_only or a.no_serve else batch_leg(batch_problems(), modes=tuple(a.batch_modes.split(",")),
                                                             tune=a.tune)
    estimation = None if a.main_only or a.no_serve else mce_leg(a.mce_images)
    config1 = None if a.main_only else config1_leg()
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(p, eps=results[a.precision]["eps"])
        log(f"[bench] cpu baseline {cpu['value']:.3f} grad evals/s")
    r = results[a.precision]
    line = {
        "metric": METRIC,
        "value": round(r["grad_evals_s"], 2),
        "unit": "grad evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["ms_per_iter"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if a.precision == "fp32" else "f64",
        "data": "synthetic (seeded GP draw + noise; W from the device kriging build)",
        "config": {
            "workload": "dual-res 1mm: highres 128x160x128 @1mm + stdres 64x80x64 @2mm, embedding 256x320x256, "
                        "L=25, thin 3, neighbourhood 6.9mm, tau2=0.806 psi=0.132 nu=1.5, state mode compat",
            "n_h": int(p["Yh"].size), "n_s": int(p["Ys"].size), "W_nnz": int(p["W"].nnz),
            "kriging_unique_solves": int(p["kriging_unique"]), "kriging_build_s": round(p["kriging_s"], 2),
            "W_layout": wst,
        },
        "fp64_value": round(results["fp64"]["grad_evals_s"], 2) if "fp64" in results else None,
        # same seed: the FP32 chain's posterior mean over its saved samples
        # against the FP64 chain's, and that difference over the Monte-Carlo
        # spread of two independent chains' means (~1: all Monte-Carlo error)
        "fp32_vs_fp64_posterior_mean_rel": None if mean_rel is None else round(mean_rel, 8),
        "fp32_vs_fp64_mean_diff_over_mc_spread": None if mc_ratio is None else round(mc_ratio, 4),
        "fp32_real_state_value": round(rr["grad_evals_s"], 2) if rr else None,
        "captured_iterations": {"value": round(rc["grad_evals_s"], 2), "ms_per_step": round(rc["ms_per_iter"], 4),
                                "p50_ms": round(rc["p50_ms"], 3), "p99_ms": round(rc["p99_ms"], 3),
                                "bitwise_equal_to_eager": rc["bitwise_equal_to_eager"]},
        "iteration_latency_ms": {"p50": round(r["p50_ms"], 3), "p99": round(r["p99_ms"], 3)},
        "leapfrog_microbench": lf,
        "kernel_ms_per_leapfrog_step": {k: round(v, 5) for k, v in ks.items()},
        "fft_passes_hbm_gbs": round(fft_bytes / (fft_ms * 1e-3) / 1e9, 1),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(dom, a.precision),
                     "algorithmic_bytes_per_launch": int(byt), "launch_ms": round(ks[dom], 5)},
        "serving_latency": serving,
        "batch_throughput": batch,
        "parameter_estimation": estimation,
        "config1_fit": config1,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
