"""Learned limiter: training data made by the method itself, training on the
GPU, and the evaluation on the triple point.

Data.  Four-quadrant 2D Riemann problems on [0, 1]^2 (quadrant states drawn
as in the batching form: density log-uniform in [0.1, 10], pressure
log-uniform in [0.01, 10], x and y velocity uniform in [-1, 1]; split point
uniform in [0.3, 0.7]^2; one perfect gas with gamma uniform in [1.2, 1.67];
reflecting walls; final time uniform in [0.05, 0.15]) run with the plain
method (minmod) on fine meshes as one batch, then restricted 4x: coarse
density = mean of the 16 fine densities (the cells have equal volumes after
the Euler remap), coarse specific energy = mass-weighted mean, coarse node
velocity = the fine node at the same position.  The network inputs of the
coarse state come from the library's own feature kernel (the one inference
runs).  The target of a coarse cell, per direction d and channel k (density,
energy, cell-mean x and y velocity): the central slope of the coarse data
(x_{+1} - x_{-1}) / 2 and the least-squares slope of the 16 fine values of the
cell per coarse cell width, both divided by the channel's feature scale.  The
loss (alpha_d sc - st)^2 makes alpha_d the factor that turns the coarse
central slope into the sub-cell slope the refined solution has.

Cells whose stencil carries no structure (relative variation below 1e-4 in
every channel) are sampled at a tenth of the rate of the others; at most
`per_problem` samples are kept per problem.

With materials > 1 the quadrants belong to 2 or 3 perfect gases (each with
its own gamma, uniform in [1.2, 1.67]; at least two different gases per
problem), so the data holds material interfaces: the coarse volume fractions
are the means of the fine ones, and the coarse state, with its membership
masks and fractions, is uploaded to a context of as many environments before
the feature kernel runs.
"""
import time

import numpy as np

from . import abi, cases

QUAD_SEED = 20261205
REFINE = 4


def quadrant_specs(n, seed=QUAD_SEED, materials=1):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        q = []
        for _ in range(4):
            q.append(dict(rho=float(np.exp(rng.uniform(np.log(0.1), np.log(10.0)))),
                          p=float(np.exp(rng.uniform(np.log(0.01), np.log(10.0)))),
                          u=float(rng.uniform(-1.0, 1.0)), v=float(rng.uniform(-1.0, 1.0))))
        out.append(dict(quads=q, x0=float(rng.uniform(0.3, 0.7)), y0=float(rng.uniform(0.3, 0.7)),
                        gamma=float(rng.uniform(1.2, 1.67)), t_end=float(rng.uniform(0.05, 0.15))))
    if materials > 1:  # gases of the quadrants from a second stream: the draws above are unchanged
        rng2 = np.random.default_rng(seed + 7)
        for sp in out:
            env = rng2.integers(0, materials, 4)
            while len(set(env.tolist())) < 2:
                env = rng2.integers(0, materials, 4)
            sp["env"] = [int(k) for k in env]
            sp["gammas"] = [float(g) for g in rng2.uniform(1.2, 1.67, materials)]
    return out


def _quadrant(spec, x, y):
    """Quadrant index (0: x<x0,y<y0; 1: x>=x0,y<y0; 2: x<x0,y>=y0; 3: x>=x0,y>=y0)."""
    return (x >= spec["x0"]).astype(int) + 2 * (y >= spec["y0"]).astype(int)


def quadrant_fields(spec, n):
    xc = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xc, xc)
    qc = _quadrant(spec, X, Y).ravel()
    rho = np.array([q["rho"] for q in spec["quads"]])[qc]
    p = np.array([q["p"] for q in spec["quads"]])[qc]
    xn = np.arange(n + 1) / n
    XN, YN = np.meshgrid(xn, xn)
    qn = _quadrant(spec, XN, YN).ravel()
    vel = np.stack([np.array([q["u"] for q in spec["quads"]])[qn], np.array([q["v"] for q in spec["quads"]])[qn]])
    env = np.array(spec.get("env", [0, 0, 0, 0]), dtype=np.int32)[qc]
    return env, rho, p, vel


def spec_gammas(spec):
    return spec.get("gammas", [spec["gamma"]])


def quadrant_mesh_options(lib, spec, n, **optkw):
    opt = lib.options(**optkw)
    abi.set_bcs(opt, abi.reflecting_walls(2))
    abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, g, 0.) for g in spec_gammas(spec)])
    opt.final_time = spec["t_end"]
    return abi.mesh_desc(2, [n, n], [0, 0], [1.0, 1.0]), opt


def restrict(n, rho, e, vel, r=REFINE):
    """Fine n x n state -> coarse (n/r)^2: density mean, energy mass-weighted
    mean, node velocity injected."""
    m = n // r
    R = rho.reshape(m, r, m, r)
    E = e.reshape(m, r, m, r)
    rc = R.mean(axis=(1, 3))
    ec = (R * E).sum(axis=(1, 3)) / R.sum(axis=(1, 3))
    V = vel.reshape(2, n + 1, n + 1)[:, ::r, ::r]
    return rc.ravel(), ec.ravel(), V.reshape(2, -1)


def restrict_fractions(n, frac, r=REFINE):
    """Fine volume fractions [ne][n*n] -> coarse [ne][(n/r)^2] (block means)."""
    m = n // r
    f = frac.reshape(-1, m, r, m, r)
    return f.mean(axis=(2, 4)).reshape(f.shape[0], -1)


def fine_slopes(n, rho, e, vel, r=REFINE):
    """Least-squares x and y slopes (per coarse cell width) of the r x r fine
    values inside each coarse cell, for density, energy and the fine
    cell-mean velocities: [2 (d), 4 (k), ncoarse]."""
    m = n // r
    u = vel.reshape(2, n + 1, n + 1)
    uc = 0.25 * (((u[:, :-1, :-1] + u[:, :-1, 1:]) + u[:, 1:, 1:]) + u[:, 1:, :-1])
    chans = [rho.reshape(n, n), e.reshape(n, n), uc[0], uc[1]]
    xi = (np.arange(r) - (r - 1) / 2) / r   # fine-cell offsets in coarse widths
    den = r * np.sum(xi * xi)
    out = np.zeros((2, 4, m * m))
    for k, ch in enumerate(chans):
        B = ch.reshape(m, r, m, r)  # [J, a (y), I, b (x)]
        out[0, k] = (np.einsum("JaIb,b->JI", B, xi) / den).ravel()
        out[1, k] = (np.einsum("JaIb,a->JI", B, xi) / den).ravel()
    return out


def samples_from_coarse(lib, m, rc, ec, vc, slopes, gammas, rng, per_problem, frac=None):
    """Features of the coarse state (GPU kernel) and targets of the kept cells;
    frac: coarse volume fractions [ne][m*m] of a multi-material problem."""
    opt = lib.options()
    abi.set_bcs(opt, abi.reflecting_walls(2))
    abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, g, 0.) for g in gammas])
    s = abi.Simulation(lib, abi.mesh_desc(2, [m, m], [0, 0], [1.0, 1.0]), opt)
    s.init_cells(np.zeros(m * m, dtype=np.int32), rc, (gammas[0] - 1.0) * rc * ec, vc)
    if frac is not None:  # membership and fractions (the features read density, energy, u, mask, fractions)
        mask = np.zeros(m * m, dtype=np.int32)
        for k in range(frac.shape[0]):
            mask |= (frac[k] > 0).astype(np.int32) << k
        s.upload({"density": np.ascontiguousarray(rc), "internal_energy": np.ascontiguousarray(ec),
                  "env_mask": mask, "fracvol_env": np.ascontiguousarray(frac.ravel())})
    X, S = s.learned_features()
    s.close()
    # central slopes from the features (x: cells (i+1, j), (i-1, j); y: (i, j+1), (i, j-1))
    sc = np.zeros((2, 4, m * m), dtype=np.float32)
    for k in range(4):
        sc[0, k] = 0.5 * (X[:, k * 25 + 13] - X[:, k * 25 + 11])
        sc[1, k] = 0.5 * (X[:, k * 25 + 17] - X[:, k * 25 + 7])
    st = (slopes / S[:, :4].T[None, :, :]).astype(np.float32)
    uc = 0.25 * (vc.reshape(2, m + 1, m + 1)[:, :-1, :-1] + vc.reshape(2, m + 1, m + 1)[:, :-1, 1:]
                 + vc.reshape(2, m + 1, m + 1)[:, 1:, 1:] + vc.reshape(2, m + 1, m + 1)[:, 1:, :-1])
    ctr = np.stack([rc, ec, uc[0].ravel(), uc[1].ravel()])
    floor = np.array([0., 0., 1.0, 1.0])[:, None]
    rel = (S[:, :4].T - 1e-6 * np.abs(ctr)) / np.maximum(np.abs(ctr), floor)
    active = np.max(rel, axis=0) > 1e-4
    keep = active | (rng.random(m * m) < 0.1 * np.mean(active))
    idx = np.flatnonzero(keep)
    if idx.size > per_problem:
        idx = rng.choice(idx, per_problem, replace=False)
    T = np.stack([sc, st], axis=-1)  # [2, 4, nc, 2]
    T = np.transpose(T, (2, 0, 1, 3)).reshape(m * m, 16)
    return X[idx], T[idx]


def make_dataset(lib, n_problems, fine=512, chunk=256, per_problem=512, seed=QUAD_SEED,
                 precision=abi.PREC_FP32_BULK, materials=1, log=print):
    """Runs the seeded four-quadrant problems (in batches of `chunk`) and
    returns the training samples (X [m][128], T [m][16])."""
    specs = quadrant_specs(n_problems, seed, materials)
    rng = np.random.default_rng(seed + 1)
    Xs, Ts = [], []
    zc_total, t_run = 0, 0.0
    for c0 in range(0, n_problems, chunk):
        sub = specs[c0:c0 + chunk]
        mo = [quadrant_mesh_options(lib, sp, fine, precision=precision) for sp in sub]
        b = abi.Batch(lib, [m for m, _ in mo], [o for _, o in mo])
        for i, sp in enumerate(sub):
            b.problem(i).init_cells(*quadrant_fields(sp, fine))
        t0 = time.perf_counter()
        nrun, zc = b.run()
        t_run += time.perf_counter() - t0
        zc_total += zc
        for i, sp in enumerate(sub):
            st = b.problem(i).download(["density", "internal_energy", "velocity", "fracvol_env"])
            rho, e, vel = st["density"], st["internal_energy"], st["velocity"]
            if not (np.all(np.isfinite(rho)) and np.all(np.isfinite(e)) and rho.min() > 0):
                continue
            rc, ec, vc = restrict(fine, rho, e, vel)
            fc = restrict_fractions(fine, st["fracvol_env"]) if materials > 1 else None
            X, T = samples_from_coarse(lib, fine // REFINE, rc, ec, vc, fine_slopes(fine, rho, e, vel),
                                       spec_gammas(sp), rng, per_problem, fc)
            Xs.append(X)
            Ts.append(T)
        b.close()
        log(f"  problems {c0 + len(sub)}/{n_problems}: {sum(len(x) for x in Xs)} samples, "
            f"{zc_total:.3e} zone-cycles in {t_run:.1f} s")
    return np.concatenate(Xs), np.concatenate(Ts), {"zone_cycles": zc_total, "run_s": t_run}


def init_params(seed):
    """He-normal initial weights; output bias 1 (the central slope)."""
    rng = np.random.default_rng(seed)
    parts = [rng.normal(0, np.sqrt(2 / 125), (64, 128)), np.zeros(64), rng.normal(0, np.sqrt(2 / 64), (64, 64)),
             np.zeros(64), rng.normal(0, 0.01, (2, 64)), np.ones(2)]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def train(lib, X, T, epochs=20, batch=1024, lr=1e-3, seed=0, params=None, lr_end=None, log=None):
    """Adam on the GPU, one epoch per call so that progress is reported; the
    optimiser state carries over between the calls.  With lr_end the rate
    follows a cosine from lr (first epoch) to lr_end (last)."""
    net = abi.Mlp(lib, init_params(seed) if params is None else params)
    net.set_data(X, T)
    losses = []
    for ep in range(epochs):
        r = lr if lr_end is None else lr_end + 0.5 * (lr - lr_end) * (1 + np.cos(np.pi * ep / max(1, epochs - 1)))
        losses.append(float(net.train(1, batch, r, seed + 1000 * ep)[0]))
        if log:
            log(f"  epoch {ep + 1}/{epochs}: loss {losses[-1]:.5f}")
    p = net.params()
    net.close()
    return p, np.array(losses)


def minmod_loss(X, T):
    """Loss of the minmod slope (the plain method) on the same targets."""
    t = T.reshape(-1, 2, 4, 2)
    out = 0.0
    for d, (lo, hi) in enumerate([(11, 13), (7, 17)]):
        for k in range(4):
            gb = -X[:, k * 25 + lo].astype(np.float64)
            gf = X[:, k * 25 + hi].astype(np.float64)
            mm = np.where(gb * gf > 0, np.sign(gb) * np.minimum(np.abs(gb), np.abs(gf)), 0.0)
            out += np.sum((mm - t[:, d, k, 1]) ** 2)
    return float(out / t.shape[0])


def baseline_loss(T, alpha):
    """Loss of a constant coefficient (1: the central slope)."""
    t = T.reshape(-1, 2, 4, 2)
    e = alpha * t[..., 0] - t[..., 1]
    return float(np.sum(e * e) / t.shape[0])


# ---- evaluation on the triple point -------------------------------------------
def triple_point_to(lib, nx, ny, t_end, **optkw):
    opt = lib.options(**optkw)
    abi.set_bcs(opt, abi.reflecting_walls(2))
    abi.set_envs(opt, [(abi.EOS_PERFECT_GAS, g, 0.) for g in cases.TP_GAMMA])
    opt.final_time = t_end
    opt.deltat_max = 1e30
    opt.land_on_final_time = 1  # every resolution ends at t_end exactly
    s = abi.Simulation(lib, abi.mesh_desc(2, [nx, ny], [0, 0], list(cases.TP_DOMAIN)), opt)
    s.init_triple_point(cases.TP_RHO, cases.TP_P)
    return s


def block_mean(a, nx, ny, r):
    return a.reshape(ny // r, r, nx // r, r).mean(axis=(1, 3)).ravel()


def evaluate_tp(lib, params, t_end=1.0, base=(896, 384), factors=(1.0, 1.5, 2.0), ref_factor=4,
                iface_alphas=(), log=print):
    """Relative L1 density error on the base grid against a plain run at
    ref_factor x the base resolution (block means): plain runs at base x
    factors and the learned run at the base resolution (material interfaces
    on minmod, network everywhere, and a fixed interface coefficient for each
    of iface_alphas)."""
    def run(nx, ny, net=None, mode=abi.LIM_LEARNED, iface=-1.0):
        kw = {"projection_limiteur_id": mode} if net is not None else {}
        s = triple_point_to(lib, nx, ny, t_end, **kw)
        if net is not None:
            s.set_limiter_mlp(net)
            if iface >= 0:
                s.set_learned_interface_alpha(iface)
        t0 = time.perf_counter()
        steps = s.run(1 << 30)
        el = time.perf_counter() - t0
        st = s.download(["density"])
        s.close()
        return st["density"], st["global_time"], steps, el

    bx, by = base
    rx, ry = bx * ref_factor, by * ref_factor
    ref, t_ref, n_ref, el_ref = run(rx, ry)
    log(f"  reference {rx}x{ry}: {n_ref} steps to t = {t_ref:.5f} in {el_ref:.1f} s")
    out = {"t_end": t_end, "reference": {"mesh": [rx, ry], "steps": n_ref, "time": t_ref}, "plain": {}}
    for f in factors:
        nx, ny = int(round(bx * f)), int(round(by * f))
        rho, t, n, el = run(nx, ny)
        # common grid: the base mesh (block means of every solution); for a
        # factor 1.5 mesh, first refine 2x by repetition then average 3x3
        if abs(f - round(f)) < 1e-9:
            rb = block_mean(rho, nx, ny, int(round(f)))
        else:
            k = int(round(2 * f))
            rr = np.repeat(np.repeat(rho.reshape(ny, nx), 2, axis=0), 2, axis=1)
            rb = block_mean(rr.ravel(), 2 * nx, 2 * ny, k)
        refb = block_mean(ref, rx, ry, ref_factor)
        err = float(np.sum(np.abs(rb - refb)) / np.sum(np.abs(refb)))
        out["plain"][f"{nx}x{ny}"] = {"l1": err, "steps": n, "time": t, "s": round(el, 2)}
        log(f"  plain {nx}x{ny}: L1 {err:.4e} ({n} steps, t = {t:.5f})")
    net = abi.Mlp(lib, params)
    s = triple_point_to(lib, bx, by, t_end, projection_limiteur_id=abi.LIM_LEARNED)
    s.set_limiter_mlp(net)
    a0 = s.learned_alpha()
    s.run(max(1, out["plain"][f"{bx}x{by}"]["steps"] // 2))
    a1 = s.learned_alpha()
    s.close()
    out["alpha_stats"] = {"start": [float(np.mean(a0)), float(np.min(a0)), float(np.max(a0))],
                          "midway": [float(np.mean(a1)), float(np.percentile(a1, 5)), float(np.percentile(a1, 95))]}
    log(f"  alpha mean/p5/p95 midway: {out['alpha_stats']['midway']}")
    refb = block_mean(ref, rx, ry, ref_factor)
    out["learned"] = {}
    variants = [("interfaces_minmod", abi.LIM_LEARNED, -1.0), ("everywhere", abi.LIM_LEARNED_ALL, -1.0)]
    variants += [(f"interfaces_alpha_{a:g}", abi.LIM_LEARNED, a) for a in iface_alphas]
    for name, mode, iface in variants:
        rho, t, n, el = run(bx, by, net, mode, iface)
        err = float(np.sum(np.abs(rho - refb)) / np.sum(np.abs(refb)))
        out["learned"][name] = {"mesh": [bx, by], "l1": err, "steps": n, "time": t, "s": round(el, 2)}
        log(f"  learned ({name}) {bx}x{by}: L1 {err:.4e} ({n} steps, t = {t:.5f})")
    net.close()
    # plain-method resolution factor with the same error (power law through
    # the base and 1.5x plain errors)
    e0 = out["plain"][f"{bx}x{by}"]["l1"]
    f15 = [k for k in out["plain"] if k != f"{bx}x{by}"][0]
    e1 = out["plain"][f15]["l1"]
    order = np.log(e0 / e1) / np.log(1.5)
    out["convergence_order_plain"] = float(order)
    for v in out["learned"].values():
        v["equivalent_plain_resolution_factor"] = float((e0 / v["l1"]) ** (1.0 / order))
    return out


def main(argv=None):
    import argparse
    import json
    import os

    from . import load_gpu
    ap = argparse.ArgumentParser(description="train the learned limiter on the GPU and evaluate it")
    ap.add_argument("--problems", type=int, default=4096)
    ap.add_argument("--fine", type=int, default=512)
    ap.add_argument("--chunk", type=int, default=256)
    ap.add_argument("--per-problem", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--lr-end", type=float, default=None, help="cosine decay of the rate down to this value")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                  "limiter_mlp.npz"))
    ap.add_argument("--materials", type=int, default=1, help="gases per four-quadrant problem (1-3)")
    ap.add_argument("--eval-t", type=float, default=1.0)
    ap.add_argument("--eval-base", type=int, nargs=2, default=[896, 384])
    ap.add_argument("--no-eval", action="store_true")
    ap.add_argument("--eval-only", metavar="NPZ", help="evaluate saved parameters, no training")
    a = ap.parse_args(argv)
    lib = load_gpu()
    if a.eval_only:
        params = np.load(a.eval_only)["params"]
        ev = evaluate_tp(lib, params, a.eval_t, base=tuple(a.eval_base), log=lambda s: print(s, flush=True))
        print(json.dumps(ev))
        return 0
    t0 = time.perf_counter()
    print(f"data: {a.problems} four-quadrant problems at {a.fine}^2, restricted {REFINE}x", flush=True)
    X, T, info = make_dataset(lib, a.problems, a.fine, a.chunk, a.per_problem, materials=a.materials,
                              log=lambda s: print(s, flush=True))
    t_data = time.perf_counter() - t0
    nval = max(1, len(X) // 20)
    Xtr, Ttr, Xva, Tva = X[nval:], T[nval:], X[:nval], T[:nval]
    print(f"samples: {len(Xtr)} train, {nval} validation ({t_data:.1f} s); "
          f"central-slope loss {baseline_loss(Ttr, 1.0):.4f}, first-order loss {baseline_loss(Ttr, 0.0):.4f}",
          flush=True)
    t0 = time.perf_counter()
    params, losses = train(lib, Xtr, Ttr, a.epochs, a.batch, a.lr, lr_end=a.lr_end,
                           log=lambda s: print(s, flush=True) if a.epochs <= 100 else None)
    t_train = time.perf_counter() - t0
    net = abi.Mlp(lib, params)
    net.set_data(Xva, Tva)
    val_loss, _ = net.gradient(0, len(Xva))
    net.close()
    print(f"training: {a.epochs} epochs in {t_train:.1f} s, loss {losses[0]:.4f} -> {losses[-1]:.4f}, "
          f"validation {val_loss:.4f} (central {baseline_loss(Tva, 1.0):.4f}, minmod {minmod_loss(Xva, Tva):.4f})",
          flush=True)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    summary = {"problems": a.problems, "materials": a.materials, "fine": a.fine, "samples": int(len(Xtr)),
               "epochs": a.epochs,
               "batch": a.batch, "lr": a.lr, "lr_end": a.lr_end, "epoch_loss": [float(x) for x in losses],
               "validation_loss": float(val_loss), "validation_loss_central": baseline_loss(Tva, 1.0),
               "validation_loss_first_order": baseline_loss(Tva, 0.0),
               "validation_loss_minmod": minmod_loss(Xva, Tva), "data_s": round(t_data, 1),
               "train_s": round(t_train, 1), "fine_zone_cycles": int(info["zone_cycles"]),
               "fine_run_s": round(info["run_s"], 1)}
    if not a.no_eval:
        summary["evaluation"] = evaluate_tp(lib, params, a.eval_t, base=tuple(a.eval_base),
                                            log=lambda s: print(s, flush=True))
    np.savez(a.out, params=params, summary=json.dumps(summary))
    print(json.dumps(summary))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
