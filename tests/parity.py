"""Helpers shared by the GPU parity tests: run the same setup through the
HIP library and the oracle and compare."""
import numpy as np

from oracle import oracle
from phantom_amd import sph


def run_gpu(setup, icall=1):
    p = setup["params"]
    ctx = sph.Context(p)
    xyzh = setup["xyzh"].copy()
    vxyzu = setup["vxyzu"]
    alphaind = setup["alphaind"].copy()
    iphase = setup.get("iphase")
    leaf, ncells = ctx.maketree(xyzh, iphase=iphase)
    tree = ctx.get_tree()
    d = ctx.densityiterate(icall, xyzh, vxyzu, Bevol=setup.get("Bevol"), alphaind=alphaind, iphase=iphase)
    stats = ctx.get_neighbour_stats()
    nneigh, nits = ctx.get_particle_stats(xyzh.shape[0])
    c = ctx.cons2prim(xyzh, vxyzu, d["dvdx"], alphaind, dustevol=setup.get("dustevol"))
    bins = setup.get("bins")
    if bins is not None:
        ibin = bins["ibin"].copy()
        ibin_wake = bins["ibin_wake"].copy()
        ctx.set_bins(ibin, bins["ibin_old"], ibin_wake, bins["nbinmax"], bins["istepfrac"], bins["ibinnow"])
    f = ctx.force(1, xyzh, vxyzu, d["gradh"], alphaind, c["eos_vars"], divcurlB=d["divcurlB"],
                  Bevol=setup.get("Bevol"), dustfrac=c["dustfrac"], iphase=iphase)
    if bins is not None:
        f.update(ibin=ibin, ibin_wake=ibin_wake, nbinmax=ctx.get_nbinmax())
    out = dict(ctx=ctx, xyzh=xyzh, leaf=leaf, ncells=ncells, tree=tree, dens=d, stats=stats, nneigh=nneigh,
               nits=nits, c2p=c, force=f, alphaind=alphaind)
    return out


def run_oracle(setup, icall=1):
    return oracle.run_all(setup, icall=icall)


def rel_rms(a, b):
    """max |a - b| / rms(b) (a normalised error that stays meaningful where
    b crosses zero)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.sqrt(np.mean(b * b))
    if scale == 0.0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)


def rel_err(a, b):
    """max over particles of |a - b| / max(|b|, rms(b)): the per-particle
    relative error, measured against the rms where b passes through zero."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.maximum(np.abs(b), np.sqrt(np.mean(b * b)))
    scale = np.where(scale > 0.0, scale, 1.0)
    return float(np.max(np.abs(a - b) / scale))


def reachable_nodes(tree):
    """Indices (0-based) of the nodes reachable from the root."""
    node = tree["node"]
    leaf = tree.get("leaf_is_active")
    n = len(node)
    used = np.zeros(n, dtype=bool)
    used[0] = True
    for k in range(n):
        if not used[k]:
            continue
        l, r = node["leftchild"][k], node["rightchild"][k]
        if l > 0 and (leaf is None or leaf[k] == 0):
            used[l - 1] = True
            used[r - 1] = True
    return np.nonzero(used)[0]
