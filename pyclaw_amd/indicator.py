"""Learned troubled-interface indicator for SharpClaw's WENO5 path.

A 6 -> 32 -> 32 -> 1 MLP (BF16 MFMA, FP32 accumulation; pyclaw_amd/csrc/
indicator.hip) looks at the min-max normalised 6-cell WENO5 stencil of every
(interface, conserved component) and predicts whether the WENO5 weights of that
component depart from the linear weights.  Interfaces with no flagged component
(and no flagged neighbour along the normal) are reconstructed with WENO5's
linear 5th-order weights; flagged ones keep the full characteristic WENO5 of
the reference (sharpclaw/reconstruct.f90:201-299), which stays the judge of
every flagged interface.  There is no counterpart in the reference.

Training runs on the GPU (clawhip_mlp_train_step): labels come from full-WENO
SharpClaw runs of seeded shallow-water problems (generate_training_data).
"""
import os

import numpy as np
import torch

from . import kernels as K

DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "indicator_swe.npy")


def init_params(seed=0):
    """He-style initialisation in the kernel's parameter layout; W1 rows 7..15
    (padding) and the biases start at zero."""
    nf, nh = 16, 32
    rng = np.random.default_rng(seed)
    p = np.zeros(K.mlp_param_count(), np.float32)
    w1 = rng.normal(0, np.sqrt(2.0 / 7.0), (nf, nh)).astype(np.float32)
    w1[6:] = 0.0
    p[:nf * nh] = w1.ravel()
    p[nf * nh:nf * nh + nh * nh] = rng.normal(0, np.sqrt(2.0 / nh), nh * nh)
    o = nf * nh + nh * nh + nh
    p[o:o + nh] = rng.normal(0, np.sqrt(1.0 / nh), nh)
    return p


class TroubledInterfaceIndicator(object):
    """Per-interface WENO5 selector.

    params: FP32 parameter vector (clawhip_mlp_param_count floats) or None for
    the shipped shallow-water weights.  threshold: logit above which a component
    is troubled; flat: stencils whose range is at most `flat` are never troubled
    (None: the values stored with the shipped weights).
    """

    def __init__(self, params=None, threshold=None, flat=None, device="cuda"):
        meta = {}
        if params is None:
            params, meta = load(DEFAULT_WEIGHTS)
        self.params = torch.as_tensor(np.asarray(params, np.float32), device=device).contiguous()
        self.threshold = float(meta.get("threshold", 0.0) if threshold is None else threshold)
        self.flat = float(meta.get("flat", 1e-6) if flat is None else flat)
        self._flags = None
        self._list = None       # (ids, count) of the interfaces the stages reconstruct in full
        self._recon = {}        # dtype -> work buffer of their full-WENO5 edge values
        self.every = None       # steps between flag refreshes; None: from the CFL bound
        self._age = None
        # full WENO5 judges the indicator's steps (SharpClawSolver._judge): a step
        # that loses positivity is recomputed with full WENO5, and every
        # audit_every steps the step is compared with the full-WENO5 one
        self.audit_every = 100
        self.audit_tol = 1e-4   # relative L1 between the two states, per step
        self.judge_stats = {"steps": 0, "positivity_fallbacks": 0, "audits": 0,
                            "audit_failures": 0, "last_audit_rel_l1": None,
                            "max_audit_rel_l1": 0.0}

    def flags(self, q, bc_lower, bc_upper, reflect=(1, 2)):
        """uint8 flags of every interface of the planar state q (meqn, my, mx)."""
        meqn, my, mx = q.shape
        n = K.mlp_num_interfaces(mx, my)
        if self._flags is None or self._flags.numel() != n or self._flags.device != q.device:
            self._flags = torch.empty(n, dtype=torch.uint8, device=q.device)
        K.mlp_flags(q, bc_lower, bc_upper, reflect, self.flat, self.params, self.threshold,
                    self._flags)
        self._list = K.mlp_flag_list(self._flags, mx, my, self._list)
        return self._flags

    @property
    def flag_list(self):
        """(ids, count) of the last flags() call: the interfaces reconstructed with
        the full WENO5."""
        return self._list

    def recon_buffer(self, like):
        """The work buffer of the full-WENO5 edge values for states like `like`
        (planar (meqn, my, mx))."""
        meqn, my, mx = like.shape
        n = K.recon_elems(mx, my, meqn)
        key = (like.dtype, str(like.device))
        b = self._recon.get(key)
        if b is None or b.numel() < n:
            b = torch.empty(n, dtype=like.dtype, device=like.device)
            self._recon[key] = b
        return b

    def refresh_interval(self, cfl_max=None):
        """Steps between flag refreshes: `every` when set, else 2 when the
        solver's CFL bound is at most 0.5 (a wave then moves at most one cell in
        two steps, which the stage kernel's one-interface dilation of the flags
        covers), else 1."""
        if self.every is not None:
            return int(self.every)
        return 2 if cfl_max is not None and cfl_max <= 0.5 else 1

    def step_flags(self, q, bcs, reflect=(1, 2), cfl_max=None):
        """Flags for the next time step, recomputed from q every
        refresh_interval(cfl_max) steps."""
        every = self.refresh_interval(cfl_max)
        if self._age is None or self._age >= every or self._flags is None \
                or self._flags.numel() != K.mlp_num_interfaces(q.shape[2], q.shape[1]):
            self.flags(q, bcs[0], bcs[1], reflect)
            self._age = 0
        self._age += 1
        return self._flags

    def troubled_fraction(self):
        """Fraction of interfaces flagged by the last flags() call (syncs)."""
        return float(self._flags.float().mean()) if self._flags is not None else None


def save(path, params, threshold, flat):
    """Weights + (threshold, flat) as one float32 .npy (no pickle)."""
    arr = np.concatenate([np.asarray(params, np.float32), np.float32([threshold, flat])])
    np.save(path, arr, allow_pickle=False)


def load(path):
    arr = np.load(path, allow_pickle=False)
    n = arr.size - 2
    return arr[:n].astype(np.float32), {"threshold": float(arr[n]), "flat": float(arr[n + 1])}


# ---------------------------------------------------------------- training
def features(q, bc_lower, bc_upper, flat, wtol, reflect=(1, 2)):
    """(feat (n, 8) float32, label (n,) uint8) tensors of a device state."""
    meqn, my, mx = q.shape
    n = K.mlp_num_interfaces(mx, my) * meqn
    feat = torch.empty((n, 8), dtype=torch.float32, device=q.device)
    label = torch.empty(n, dtype=torch.uint8, device=q.device)
    K.mlp_features(q, bc_lower, bc_upper, reflect, flat, wtol, feat, label)
    return feat, label


def train(feat, label, params=None, steps=2000, batch=4096, lr=3e-3, pos_weight=4.0, seed=0,
          log_every=0):
    """Adam on the device; returns (params tensor, loss history)."""
    dev = feat.device
    n = label.numel()
    if params is None:
        params = init_params(seed)
    P = torch.as_tensor(np.asarray(params, np.float32), device=dev).clone()
    grad = torch.zeros_like(P)
    m1 = torch.zeros_like(P)
    m2 = torch.zeros_like(P)
    loss = torch.zeros(steps, dtype=torch.float32, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    for step in range(1, steps + 1):
        idx = torch.randint(0, n, (batch,), device=dev, dtype=torch.int32, generator=gen)
        K.mlp_train_step(feat, label, idx, P, grad, m1, m2, step, lr, pos_weight,
                         loss[step - 1:step])
        if log_every and step % log_every == 0:
            print(f"step {step} loss {loss[step - 1].item():.4f}", flush=True)
    return P, loss.cpu().numpy()


def logits(feat, params):
    n = feat.shape[0]
    out = torch.empty(n, dtype=torch.float32, device=feat.device)
    K.mlp_logits(feat, n, params, out)
    return out


def pick_threshold(z, label, recall=0.995):
    """Largest logit threshold that keeps the given recall of troubled samples."""
    zp = torch.sort(z[label.bool()]).values
    if zp.numel() == 0:
        return 0.0
    k = int(np.floor((1.0 - recall) * zp.numel()))
    return float(zp[k].item()) - 1e-3


def swe_problem(seed, mx, my):
    """Seeded shallow-water initial state on the periodic unit square (host,
    Fortran order (3, mx, my)): h = 1 + 8 Gaussian humps (amplitude U[0.05, 0.3],
    width U[0.02, 0.1]) + 2 circular dam breaks (radius U[0.05, 0.15], jump
    U[0.1, 0.5]), hu = hv = 0.  Distances are periodic."""
    rng = np.random.default_rng(seed)
    x = (np.arange(mx) + 0.5) / mx
    y = (np.arange(my) + 0.5) / my
    X, Y = np.meshgrid(x, y, indexing="ij")

    def r2(cx, cy):
        ddx = (X - cx + 0.5) % 1.0 - 0.5
        ddy = (Y - cy + 0.5) % 1.0 - 0.5
        return ddx * ddx + ddy * ddy

    h = np.ones((mx, my))
    for _ in range(8):
        amp, width = rng.uniform(0.05, 0.3), rng.uniform(0.02, 0.1)
        cx, cy = rng.uniform(0, 1, 2)
        h += amp * np.exp(-r2(cx, cy) / (width * width))
    for _ in range(2):
        rad, jump = rng.uniform(0.05, 0.15), rng.uniform(0.1, 0.5)
        cx, cy = rng.uniform(0, 1, 2)
        h += np.where(r2(cx, cy) < rad * rad, jump, 0.0)
    q = np.zeros((3, mx, my), order="F")
    q[0] = h
    return q


def swe_solver(precision="fp64", indicator=None, cfl=0.45, cfl_max=0.5):
    """SharpClawSolver2D for the periodic shallow-water family: characteristic
    WENO5, SSP33, CFL 0.45 (cfl_max 0.5, so the indicator refreshes its flags
    every second step)."""
    from . import BC, SharpClawSolver2D, riemann
    solver = SharpClawSolver2D(riemann.shallow_roe_with_efix_2D)
    solver.time_integrator = "SSP33"
    solver.cfl_desired = cfl
    solver.cfl_max = cfl_max
    solver.char_decomp = 2
    solver.all_bcs = BC.periodic
    solver.precision = precision
    solver.indicator = indicator
    return solver


def swe_solution(seed, mx, my):
    from . import Domain, Solution
    domain = Domain([0.0, 0.0], [1.0, 1.0], [mx, my])
    sol = Solution(3, domain)
    sol.problem_data["grav"] = 1.0
    sol.problem_data["efix"] = True
    sol.q = swe_problem(seed, mx, my)
    return sol


def generate_training_data(seeds, mx=512, my=512, nsteps=300, snapshots=(50, 150, 300),
                           flat=1e-6, wtol=0.05, keep_negative=0.1, device="cuda"):
    """Run full characteristic WENO5 SharpClaw (FP64, SSP33) on the seeded
    shallow-water family and collect features/labels after the given step
    counts.  All troubled samples are kept, a `keep_negative` fraction of the
    others."""
    feats, labels = [], []
    gen = torch.Generator(device=device)
    gen.manual_seed(12345)
    seeds = list(seeds)
    for k, seed in enumerate(seeds):
        if k % 16 == 0:
            print(f"training data: run {k} of {len(seeds)}", flush=True)
        solver = swe_solver("fp64", cfl=0.5, cfl_max=0.55)
        sol = swe_solution(seed, mx, my)
        solver.dt = 0.5 / mx / 1.5
        solver.setup(sol)
        for step in range(1, nsteps + 1):
            solver.evolve_to_time(sol)
            if step in snapshots:
                q, _ = sol.state.device_q(K.L.F64)
                f, l = features(q, [2, 2], [2, 2], flat, wtol)
                keep = l.bool() | (torch.rand(l.shape, device=l.device, generator=gen)
                                   < keep_negative)
                feats.append(f[keep])
                labels.append(l[keep])
    return torch.cat(feats), torch.cat(labels)


def main(argv=None):
    """python -m pyclaw_amd.indicator --out FILE: train the shipped weights."""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_WEIGHTS)
    ap.add_argument("--runs", type=int, default=256)
    ap.add_argument("--mx", type=int, default=512)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--recall", type=float, default=0.999)
    ap.add_argument("--sim-steps", type=int, default=300)
    ap.add_argument("--flat", type=float, default=1e-6)
    ap.add_argument("--wtol", type=float, default=0.05)
    ap.add_argument("--ship-flat", type=float, default=None,
                    help="the `flat` stored with the weights for inference (default: --flat)")
    args = ap.parse_args(argv)
    snaps = tuple(sorted({max(1, args.sim_steps // 6), args.sim_steps // 2, args.sim_steps}))
    feat, label = generate_training_data(range(1000, 1000 + args.runs), args.mx, args.mx,
                                         nsteps=args.sim_steps, snapshots=snaps, flat=args.flat,
                                         wtol=args.wtol)
    n = label.numel()
    perm = torch.randperm(n, device=feat.device)
    ntr = int(0.9 * n)
    tr, va = perm[:ntr], perm[ntr:]
    print(f"samples {n}, troubled fraction {label.float().mean().item():.3f}", flush=True)
    P, hist = train(feat[tr].contiguous(), label[tr].contiguous(), steps=args.steps,
                    log_every=500)
    z = logits(feat[va].contiguous(), P)
    lv = label[va]
    thr = pick_threshold(z, lv, args.recall)
    pred = z > thr
    recall = (pred & lv.bool()).sum().item() / max(1, lv.sum().item())
    flagged = pred.float().mean().item()
    print(f"threshold {thr:.3f} recall {recall:.4f} flagged {flagged:.3f} "
          f"(troubled {lv.float().mean().item():.3f})", flush=True)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    save(args.out, P.cpu().numpy(), thr,
         args.flat if args.ship_flat is None else args.ship_flat)
    print(f"saved {args.out}", flush=True)


if __name__ == "__main__":
    main()
