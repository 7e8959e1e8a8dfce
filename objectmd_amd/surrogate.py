"""Learned surrogate of the Tersoff potential: model definition, training
data from the device Tersoff kernels, training, and the parameter blocks of
omdg_set_surrogate.

The model (the same definition the device kernels evaluate):
  fc(r)   = (cos(pi r / rc) + 1) / 2 for r < rc
  G_k     = sum_j exp(-eta_r (r_ij - mu0 - k dmu)^2) fc(r_ij)           k < 24
  G_24+m  = sum_{j<k} A_m(cos t_jik) exp(-eta_a (r_ij^2 + r_ik^2))
                      fc(r_ij) fc(r_ik)                                 m < 8
  A_m(c)  = exp(-q_m (c - p_m)^2), centres p_m spread over [-1, 1] (kind 1;
            kind 0 is Behler's 2^(1-q) (1 + p c)^q)
  E_i     = w3 . silu(W2 silu(W1 (G_i - mean)/std + b1) + b2) + b3
Training data: Si slabs of 4,096 atoms (8^3 diamond cells, periodic in x
and y, free (001) surfaces) run with the device Tersoff kernels at
300-3,000 K, the hot ones melted first.  Loss on the energy per atom and on
the forces, each snapshot's force error scaled by its mean square force.  Training runs with PyTorch (autograd gives the force loss's
gradient); inference runs in the engine.
"""
import math
import os

import numpy as np

from . import POT_TERSOFF, X, F, POT, V, lattice, tables

NR, NA, ND, NH = 24, 8, 32, 64
MODEL_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                          "surrogate_si.npz")


def default_desc():
    pq = [(-1.0 + 2.0 * m / 7.0, 24.0) for m in range(NA)]
    return dict(rc=3.0, eta_r=50.0, mu0=1.5, dmu=1.5 / 23.0, eta_a=0.0, kind=1, pq=pq)


def desc_vector(d):
    out = [d["rc"], d["eta_r"], d["mu0"], d["dmu"], d["eta_a"], float(d["kind"])]
    for p, q in d["pq"]:
        out += [p, q]
    return np.array(out, np.float64)


# ---- model file ---------------------------------------------------------------
def save(path, model, desc, meta):
    np.savez(path, desc=desc_vector(desc), **{k: np.asarray(v) for k, v in model.items()},
             **{"meta_" + k: np.asarray(v) for k, v in meta.items()})


def load(path=MODEL_PATH):
    """-> (model dict, descriptor dict, metadata dict)."""
    z = np.load(path, allow_pickle=False)
    dv = z["desc"]
    desc = dict(rc=dv[0], eta_r=dv[1], mu0=dv[2], dmu=dv[3], eta_a=dv[4], kind=int(dv[5]),
                pq=[(dv[6 + 2 * m], dv[7 + 2 * m]) for m in range(NA)])
    model = {k: z[k] for k in ("W1", "b1", "W2", "b2", "w3", "b3", "mean", "std", "lo", "hi")}
    meta = {k[5:]: z[k] for k in z.files if k.startswith("meta_")}
    return model, desc, meta


def api_blocks(model, desc):
    """(desc[22], norm[128], weights[6337]) in the order of omdg_set_surrogate.
    The gauge box of a descriptor the training data left nearly constant is
    widened to 5 % of the widest one, so that it measures distance on the
    common scale rather than flagging every small excursion."""
    lo, hi = np.asarray(model["lo"], float), np.asarray(model["hi"], float)
    width = np.maximum(hi - lo, 0.05 * float((hi - lo).max()))
    mid = 0.5 * (lo + hi)
    norm = np.concatenate([model["mean"], model["std"], mid - 0.5 * width, mid + 0.5 * width])
    w = np.concatenate([np.ravel(model["W1"]), model["b1"], np.ravel(model["W2"]), model["b2"],
                        model["w3"], np.atleast_1d(model["b3"])]).astype(np.float32)
    return desc_vector(desc), norm.astype(np.float64), w


def configure(eng, model, desc, tersoff_params, threshold=0.25):
    """Tersoff parameters for the fallback steps, the model, OMDG_POT_SURROGATE."""
    from . import POT_SURROGATE
    eng.set_tersoff(tersoff_params)
    eng.set_surrogate(*api_blocks(model, desc), threshold)
    eng.set_potential(POT_SURROGATE)


# ---- training data ------------------------------------------------------------
def slab(cells=8, a=5.431):
    pos, (lo, hi) = lattice.diamond100(cells, cells, cells, a)
    return pos, lo, hi


def generate(tersoff_params, temps, per_temp, seed=7, cells=8, melt_above=2400.0, log=None):
    """Snapshots (x, forces, per-atom energies, box) of Tersoff MD on the
    device.  For each temperature: a run from the slab (from a 5,000 K melt
    when T >= melt_above), velocity rescaling every 50 steps, then
    per_temp snapshots 200 steps apart."""
    from . import Engine
    m = tables.mass_sim(tables.MASS_SI)
    pos, lo, hi = slab(cells)
    n = len(pos)
    out = []
    rng = np.random.default_rng(seed)
    for t_target in temps:
        eng = Engine()
        eng.set_box(lo, hi, 3)
        x0 = pos + rng.normal(0.0, 0.02, pos.shape)
        t_start = 5000.0 if t_target >= melt_above else t_target
        v0 = lattice.maxwell_boltzmann(n, m, 2.0 * t_start, int(rng.integers(1 << 30)))
        eng.upload_atoms(x0, v0, mass=[m])
        eng.set_tersoff(tersoff_params)
        eng.set_potential(POT_TERSOFF)
        eng.set_neighbor(0.5, 10, 0)
        eng.compute()
        eng.zero_forces()

        def run(steps, temp):
            for _ in range(steps // 50):
                eng.step(50, 0.001)
                v = eng.download(V)
                tk = lattice.kinetic_temperature(v, m)
                if tk > 0:
                    eng.set_state(v=v * math.sqrt(temp / tk))

        if t_start != t_target:
            run(1000, t_start)
        run(1500, t_target)
        for _ in range(per_temp):
            run(200, t_target)
            eng.compute()
            out.append(dict(x=eng.download(X), f=eng.download(F), e=eng.download(POT),
                            lo=lo, hi=hi, temp=t_target))
        eng.close()
        if log:
            log(f"[surrogate] {t_target:.0f} K: {per_temp} snapshots")
    return out


def pack_snapshots(snaps):
    return dict(x=np.array([s["x"] for s in snaps], np.float32),
                f=np.array([s["f"] for s in snaps], np.float32),
                e=np.array([s["e"] for s in snaps], np.float32),
                temp=np.array([s["temp"] for s in snaps]), lo=snaps[0]["lo"], hi=snaps[0]["hi"])


# ---- descriptor in PyTorch (training) -----------------------------------------
def neighbour_pairs(x, lo, hi, pbc, rc):
    """(i, j, shift) of every pair with r < rc, minimum image in the periodic
    dimensions (box > 2 rc), brute force in chunks."""
    L = hi - lo
    n = len(x)
    ii, jj, sh = [], [], []
    for a in range(0, n, 512):
        d = x[None, :, :] - x[a:a + 512, None, :]
        s = np.zeros_like(d)
        for k in range(3):
            if (pbc >> k) & 1:
                s[..., k] = -np.round(d[..., k] / L[k]) * L[k]
        d = d + s
        r2 = (d * d).sum(-1)
        ia, ja = np.nonzero((r2 < rc * rc) & (r2 > 1e-12))
        ii.append(ia + a)
        jj.append(ja)
        sh.append(s[ia, ja])
    return np.concatenate(ii), np.concatenate(jj), np.concatenate(sh)


def triplets(ii, n):
    """Ordered pairs (u, w), u != w, of pairs sharing the centre."""
    order = np.argsort(ii, kind="stable")
    start = np.searchsorted(ii[order], np.arange(n + 1))
    us, ws = [], []
    for c in range(n):
        p = order[start[c]:start[c + 1]]
        if len(p) < 2:
            continue
        u, w = np.meshgrid(p, p, indexing="ij")
        keep = u != w
        us.append(u[keep])
        ws.append(w[keep])
    if not us:
        return np.zeros(0, int), np.zeros(0, int)
    return np.concatenate(us), np.concatenate(ws)


def torch_descriptor(xt, ii, jj, sh, tu, tw, desc, n):
    import torch
    d = xt[jj] - xt[ii] + sh
    r = torch.sqrt((d * d).sum(1))
    rc = desc["rc"]
    fc = 0.5 * (torch.cos(math.pi * r / rc) + 1.0)
    mu = desc["mu0"] + desc["dmu"] * torch.arange(NR, dtype=xt.dtype, device=xt.device)
    g = torch.exp(-desc["eta_r"] * (r[:, None] - mu[None, :]) ** 2) * fc[:, None]
    G = torch.zeros((n, ND), dtype=xt.dtype, device=xt.device)
    G = G.index_add(0, ii, torch.cat([g, torch.zeros((len(r), NA), dtype=xt.dtype,
                                                     device=xt.device)], 1))
    if len(tu):
        cos = (d[tu] * d[tw]).sum(1) / (r[tu] * r[tw])
        rad = torch.exp(-desc["eta_a"] * (r[tu] ** 2 + r[tw] ** 2)) * fc[tu] * fc[tw]
        cols = []
        for p, q in desc["pq"]:
            if desc["kind"] == 1:
                a = torch.exp(-q * (cos - p) ** 2)
            else:
                a = 2.0 ** (1.0 - q) * (1.0 + p * cos) ** q
            cols.append(0.5 * a * rad)
        ang = torch.stack(cols, 1)
        G = G.index_add(0, ii[tu], torch.cat([torch.zeros((len(tu), NR), dtype=xt.dtype,
                                                          device=xt.device), ang], 1))
    return G


def train(data, tersoff_params=None, epochs=300, batch=10, lr=3e-3, seed=1, device=None,
          holdout=20, log=print):
    """Fit the MLP to the snapshots; returns (model dict, desc, meta)."""
    import torch
    torch.manual_seed(seed)
    dev = torch.device(device or ("cuda" if torch.cuda.is_available() else "cpu"))
    desc = default_desc()
    xs, fs, es = data["x"].astype(np.float64), data["f"].astype(np.float64), \
        data["e"].astype(np.float64)
    lo, hi = np.asarray(data["lo"], float), np.asarray(data["hi"], float)
    ns, n = xs.shape[0], xs.shape[1]
    rng = np.random.default_rng(seed)
    perm = rng.permutation(ns)
    test_idx, train_idx = perm[:holdout], perm[holdout:]
    structs = []
    for s in range(ns):
        ii, jj, sh = neighbour_pairs(xs[s], lo, hi, 3, desc["rc"])
        tu, tw = triplets(ii, n)
        structs.append((ii, jj, sh, tu, tw))

    def batched(idx):
        """Snapshots idx as one system: atoms and pair indices offset."""
        parts = [[], [], [], [], []]
        off_p = 0
        for k, s in enumerate(idx):
            ii, jj, sh, tu, tw = structs[s]
            parts[0].append(ii + k * n)
            parts[1].append(jj + k * n)
            parts[2].append(sh)
            parts[3].append(tu + off_p)
            parts[4].append(tw + off_p)
            off_p += len(ii)
        t = [torch.as_tensor(np.concatenate(p), device=dev) for p in parts]
        # force errors weighted by the inverse mean square force of their
        # snapshot: every temperature counts alike
        wgt = np.repeat(fscale / (fs[idx] ** 2).mean(axis=(1, 2)), n)
        return dict(x=torch.as_tensor(np.concatenate([xs[s] for s in idx]), device=dev),
                    f=torch.as_tensor(np.concatenate([fs[s] for s in idx]), device=dev),
                    w=torch.as_tensor(wgt, device=dev)[:, None],
                    e=torch.as_tensor(es[idx].sum(1), device=dev),
                    seg=torch.as_tensor(np.repeat(np.arange(len(idx)), n), device=dev),
                    st=t, m=len(idx))

    fscale = float((fs ** 2).mean())
    train_batches = [batched(train_idx[b:b + batch]) for b in range(0, len(train_idx), batch)]
    test_batches = [batched(test_idx[b:b + batch]) for b in range(0, len(test_idx), batch)]
    with torch.no_grad():
        Gs = torch.cat([torch_descriptor(tb["x"], *tb["st"], desc, tb["m"] * n)
                        for tb in train_batches])
    mean = Gs.mean(0)
    std = Gs.std(0)
    std = std.clamp_min(0.02 * float(std.max()))  # nearly constant features
    glo, ghi = Gs.min(0).values, Gs.max(0).values
    e0 = float(np.mean(es[train_idx].sum(1))) / n
    W1 = torch.nn.Parameter(torch.randn(NH, ND, dtype=torch.float64, device=dev) / math.sqrt(ND))
    b1 = torch.nn.Parameter(torch.zeros(NH, dtype=torch.float64, device=dev))
    W2 = torch.nn.Parameter(torch.randn(NH, NH, dtype=torch.float64, device=dev) / math.sqrt(NH))
    b2 = torch.nn.Parameter(torch.zeros(NH, dtype=torch.float64, device=dev))
    w3 = torch.nn.Parameter(torch.randn(NH, dtype=torch.float64, device=dev) * 0.01)
    b3 = torch.nn.Parameter(torch.tensor(e0, dtype=torch.float64, device=dev))
    params = [W1, b1, W2, b2, w3, b3]
    opt = torch.optim.Adam(params, lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, epochs)

    def energy_forces(tb, create_graph):
        x = tb["x"].clone().requires_grad_(True)
        G = torch_descriptor(x, *tb["st"], desc, tb["m"] * n)
        z = (G - mean) / std
        h1 = torch.nn.functional.silu(z @ W1.T + b1)
        h2 = torch.nn.functional.silu(h1 @ W2.T + b2)
        Ei = h2 @ w3 + b3
        f = -torch.autograd.grad(Ei.sum(), x, create_graph=create_graph)[0]
        E = torch.zeros(tb["m"], dtype=Ei.dtype, device=dev).index_add(0, tb["seg"], Ei)
        return E, f

    wf, we = 1.0, 10.0
    for ep in range(epochs):
        tot = 0.0
        for bi in rng.permutation(len(train_batches)):
            tb = train_batches[bi]
            opt.zero_grad()
            E, f = energy_forces(tb, True)
            loss = wf * (tb["w"] * (f - tb["f"]) ** 2).mean() + \
                we * (((E - tb["e"]) / n) ** 2).mean()
            loss.backward()
            opt.step()
            tot += loss.item()
        sched.step()
        if log and (ep % 10 == 0 or ep == epochs - 1):
            log(f"[surrogate] epoch {ep} loss {tot / len(train_batches):.5f}")
    # held-out errors
    fe, ee = [], []
    for tb in test_batches:
        E, f = energy_forces(tb, False)
        fe.append(((f - tb["f"]) ** 2).mean().item())
        ee.append((torch.abs(E - tb["e"]) / n).mean().item())
    model = dict(W1=W1.detach().cpu().numpy(), b1=b1.detach().cpu().numpy(),
                 W2=W2.detach().cpu().numpy(), b2=b2.detach().cpu().numpy(),
                 w3=w3.detach().cpu().numpy(), b3=np.float64(b3.detach().cpu()),
                 mean=mean.cpu().numpy(), std=std.cpu().numpy(), lo=glo.cpu().numpy(),
                 hi=ghi.cpu().numpy())
    meta = dict(force_rmse_holdout=math.sqrt(np.mean(fe)), energy_mae_holdout=float(np.mean(ee)),
                snapshots=ns, holdout=holdout, atoms=n, epochs=epochs,
                temps=np.asarray(sorted(set(np.asarray(data["temp"]).tolist()))))
    if log:
        log(f"[surrogate] hold-out force RMSE {meta['force_rmse_holdout']:.4f} eV/A, "
            f"energy MAE {meta['energy_mae_holdout'] * 1e3:.2f} meV/atom")
    return model, desc, meta


def main():
    """Generate the snapshots on the GPU, train, write the model file."""
    import argparse
    from .params import tersoff_params as read_tersoff
    ap = argparse.ArgumentParser()
    ap.add_argument("--tersoff", required=True, help="tables/tersoff.silicon")
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--data", default="", help="reuse or write the snapshots (.npz)")
    ap.add_argument("--temps", type=int, default=20)
    ap.add_argument("--per-temp", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=300)
    args = ap.parse_args()
    tp = read_tersoff(args.tersoff, 2.7, 0.15)
    log = lambda m: print(m, flush=True)
    if args.data and os.path.exists(args.data):
        data = dict(np.load(args.data, allow_pickle=False))
    else:
        temps = np.linspace(300.0, 3000.0, args.temps)
        data = pack_snapshots(generate(tp, temps, args.per_temp, log=log))
        if args.data:
            np.savez(args.data, **data)
    model, desc, meta = train(data, tp, epochs=args.epochs, log=log)
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    save(args.out, model, desc, meta)
    log(f"[surrogate] wrote {args.out}")


if __name__ == "__main__":
    main()
