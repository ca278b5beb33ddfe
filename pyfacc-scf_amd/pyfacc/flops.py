"""Algorithmic floating-point operation model of the ERI + J/K kernels.

Counts the arithmetic of the Obara-Saika VRR / HGP HRR path exactly as the
kernels in ``csrc/eri_core.h`` and ``csrc/eri.hip`` perform it, restricted to
the recurrence entries that feed the kept integrals (the compiler removes the
rest).  FMA counts 2, every other add/mul/div/sqrt/exp counts 1; loads,
integer work and screening are not counted.  This figure is the numerator of
the roofline fraction: FLOPs per launch = sum over classes of
quartets * (K_bra K_ket * per_primitive + per_contracted).
"""
from functools import lru_cache


def _nc(L):
    return (L + 1) * (L + 2) // 2


def _coff(L):
    return L * (L + 1) * (L + 2) // 6


def _comp(g):
    L = 0
    while _coff(L + 1) <= g:
        L += 1
    pos = g - _coff(L)
    r = 0
    while (r + 1) * (r + 2) // 2 <= pos:
        r += 1
    z = pos - r * (r + 1) // 2
    x = L - r
    return (x, L - x - z, z), L


def _idx(c):
    x, y, z = c
    L = x + y + z
    return _coff(L) + (y + z) * (y + z + 1) // 2 + z


def _dir(c):
    return 0 if c[0] > 0 else (1 if c[1] > 0 else 2)


def _minus(c, d):
    c = list(c)
    c[d] -= 1
    return tuple(c)


@lru_cache(None)
def per_primitive(la, lb, lc, ld, taylor_terms=7):
    """FLOPs of one primitive quartet of class (la lb|lc ld)."""
    LAB, LCD = la + lb, lc + ld
    L = LAB + LCD
    GE, GF = _coff(LAB + 1), _coff(LCD + 1)
    need = set()
    stack = [(e, f, 0) for e in range(_coff(la), GE) for f in range(_coff(lc), GF)]
    flops = {}
    while stack:
        g, h, m = stack.pop()
        if (g, h, m) in need:
            continue
        need.add((g, h, m))
        cg, Lg = _comp(g)
        ch, Lh = _comp(h)
        if Lg == 0 and Lh == 0:
            flops[(g, h, m)] = 1  # pref * F[m]
            continue
        deps = []
        if Lh == 0:  # bra step
            d = _dir(cg)
            g1 = _minus(cg, d)
            deps += [(_idx(g1), 0, m), (_idx(g1), 0, m + 1)]
            f = 3
            if g1[d] > 0:
                g2 = _minus(g1, d)
                deps += [(_idx(g2), 0, m), (_idx(g2), 0, m + 1)]
                f += 5
        else:  # ket step
            d = _dir(ch)
            h1 = _minus(ch, d)
            deps += [(g, _idx(h1), m), (g, _idx(h1), m + 1)]
            f = 3
            if h1[d] > 0:
                h2 = _minus(h1, d)
                deps += [(g, _idx(h2), m), (g, _idx(h2), m + 1)]
                f += 5
            if cg[d] > 0:
                deps.append((_idx(_minus(cg, d)), _idx(h1), m + 1))
                f += 3
        flops[(g, h, m)] = f
        stack += deps
    rec = sum(flops.values())
    nacc = (GE - _coff(la)) * (GF - _coff(lc))
    setup = 1 + 1 + 2 + 3 + 6 + 3  # p+q, 1/(p+q), rho, PQ, T, prefactor
    boys = 4 + 3 * (taylor_terms - 1) + (2 + 3 * L if L > 0 else 0)
    bra_c = 6 if LAB > 0 else 0   # WP, 1/2p, rho/p
    ket_c = 7 if LCD > 0 else 0   # WQ, 1/2q, rho/q, 1/2(p+q)
    return setup + boys + bra_c + ket_c + rec + nacc


@lru_cache(None)
def per_contracted(la, lb, lc, ld, ndens=1):
    """FLOPs per contracted quartet: HRR and the J/K digest."""
    NA, NB, NC, ND = _nc(la), _nc(lb), _nc(lc), _nc(ld)
    NF = _coff(lc + ld + 1) - _coff(lc)
    hrr = (NA * NB * NF * 2 if lb else 0) + (NA * NB * NC * ND * 2 if ld else 0)
    nab, ncd = NA * NB, NC * ND
    jdig = nab * (ncd * 2 + 2) + ncd * (nab * 2 + 1)
    kdig = ndens * 4 * (NA * NC * (NB * ND * 2 + 1))  # four blocks of equal size for s/p
    deg = 4
    return hrr + jdig + kdig + deg


CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0), (1, 1, 1, 1)]
CLASS_NAMES = ["(ss|ss)", "(ps|ss)", "(ps|ps)", "(pp|ss)", "(pp|ps)", "(pp|pp)"]


def class_flops(cls_index, quartets, kprim_pair=9, ndens=1, taylor_terms=7, orient=None, prim_quartets=None):
    """FLOPs of ``quartets`` quartets of class CLASSES[cls_index] with
    ``kprim_pair`` primitive pairs per shell pair (9 for STO-3G), or with
    ``prim_quartets`` primitive quartets in total when given (the kernels skip
    primitive pairs whose bound is negligible); ``orient`` gives
    (la, lb, lc, ld) as launched when bra and ket were swapped."""
    c = tuple(int(x) for x in orient) if orient is not None else CLASSES[cls_index]
    nprim = quartets * kprim_pair * kprim_pair if prim_quartets is None else prim_quartets
    return nprim * per_primitive(*c, taylor_terms=taylor_terms) + quartets * per_contracted(*c, ndens=ndens)


def fock_flops(class_quartets, kprim_pair=9, ndens=1, taylor_terms=7, orients=None, class_prims=None):
    return sum(class_flops(i, q, kprim_pair, ndens, taylor_terms, None if orients is None else orients[i],
                           None if class_prims is None else class_prims[i])
               for i, q in enumerate(class_quartets))
