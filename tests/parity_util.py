"""Comparisons shared by the GPU parity tests: hit lists, candidate lists
and plane cells of the HIP path against the oracle (bars stated in
test_gpu_parity.py)."""
import numpy as np

from oracle import oracle as O


def compare_hits(gh, oh, powcut):
    gk = {(h["block"], h["stage"], h["wi"], h["zi"], h["ii"]): float(h["pow"]) for h in gh}
    ok = {(h["block"], h["stage"], h["wi"], h["zi"], h["ri"]): float(h["pow"]) for h in oh}
    worst = 0.0
    for k in set(gk) | set(ok):
        if k in gk and k in ok:
            worst = max(worst, abs(gk[k] - ok[k]) / ok[k])
        else:
            v = gk.get(k, ok.get(k))
            cut = powcut[k[1]]
            assert abs(v - cut) < 1e-5 * cut, ("hit only on one side", k, v, cut)
    return worst


def compare_cands(gc, oc):
    assert len(gc) == len(oc), (len(gc), len(oc))
    for g, w in zip(gc, oc):
        assert g["numharm"] == int(w[2])
        assert abs(g["r"] - w[3]) < 1e-3
        assert abs(g["z"] - w[4]) < 0.05
        assert abs(g["w"] - w[5]) < 0.05
        assert abs(g["sigma"] - w[1]) < 1e-3


def plane_error(s, o, fft, blocks, fracs):
    """Max |gpu*norm^2 - fp64| / max(fp64, 1) over the cells the reference
    reads, for the given blocks and harmonic fractions (numharm, harmnum)."""
    banks = O.SubharmBanks(o, fp64=True)
    rstep = o.corr_uselen // 2
    lo, nd = s.windows()
    norms = s.norms()
    nfrac = (1 << (o.numharmstages - 1)) - 1
    order = [(1, 1)] + [(1 << st, k) for st in range(1, o.numharmstages)
                        for k in range(1, 1 << st, 2)]
    worst = 0.0
    for b in blocks:
        startr = o.rlo + b * rstep
        for (H, k) in fracs:
            plane, meta = O.subharm_fderivs_vol(H, k, startr, startr + rstep - 0.5, banks, o,
                                                fft, fp64=True)
            if o.inmem:
                assert (H, k) == (1, 1)
                widx = s.geom.nblocks * 0 + (len(norms) - s.geom.nblocks) + b
            else:
                widx = b * (nfrac + 1) + order.index((H, k))
            # window parity: the GPU normalised with the oracle's window
            bn = O.block_norm(o, fft, int(lo[widx]), int(nd[widx]))
            assert abs(norms[widx] - bn) <= 1e-15 * bn, (b, H, k, norms[widx], bn)
            nuse = o.corr_uselen if H == 1 else int(np.ceil(o.corr_uselen * k / H)) + 1
            nuse = min(nuse, meta["numrs"])
            hb0 = 2 * meta["rlo"]
            for wi in range(meta["numws"]):
                wrow = (meta["wlo"] + 20 * wi - int(o.wlo)) // 20 if o.numw else 0
                for zi in range(meta["numzs"]):
                    z = meta["zlo"] + 2 * zi
                    row = (z - int(o.zlo)) // 2
                    raw = s.plane_row(wrow, row, hb0, nuse).astype(np.float64)
                    gpu = raw * np.float32(norms[widx] ** 2)
                    ref = plane[wi, zi, :nuse]
                    worst = max(worst,
                                float(np.max(np.abs(gpu - ref) / np.maximum(ref, 1.0))))
    banks.close()
    return worst
