"""Exact rational evaluation of the GPP sigma band loop for tiny shapes.

Every float64 input is a dyadic rational, so the mathematical value of the
reference's loop (gppComplex.cpp:16-150, :335-383) can be computed with no
rounding at all using fractions.Fraction.  Branches of ssxt_scht_solver are
decided on exact values; thresholds are the doubles the reference compares
against (1e-6, 0.25, 1.0/(1e-6*4.0)).  Used to check the oracle's
arithmetic (indices, signs, conjugations, term order) independently of its
floating-point rounding.
"""
from __future__ import annotations

from fractions import Fraction as Fr

TO1 = Fr(1e-6)
LIMITTWO = Fr(1, 4)
LIMITONE = Fr(1.0 / (1e-6 * 4.0))


def _c(z):
    return (Fr(float(z.real)), Fr(float(z.imag)))


def _mul(a, b):
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1])


def _conj(a):
    return (a[0], -a[1])


def _scale(a, s):
    return (a[0] * s, a[1] * s)


def _abs2(a):
    return a[0] * a[0] + a[1] * a[1]


def _mag(a):
    return abs(float(a[0])) + abs(float(a[1]))


def pole_terms_exact(wx, wt, ie):
    """(branch, sch, ssx) of ssxt_scht_solver in exact arithmetic."""
    wt2 = _mul(wt, wt)
    om2 = _mul(wt2, ie)
    wdiff = (wx - wt[0], -wt[1])
    wdiffr = _abs2(wdiff)
    if wdiffr == 0:
        raise ZeroDivisionError("wtilde == wx")
    delw = _scale(_mul(wt, _conj(wdiff)), 1 / wdiffr)
    delwr = _abs2(delw)
    zero = (Fr(0), Fr(0))
    if wdiffr > LIMITTWO and delwr < LIMITONE:
        br, sch, ssx = 1, _mul(delw, ie), _scale(om2, 1 / (wx * wx))
    elif delwr > TO1:
        cden = _scale(_mul(wt2, (Fr(1, 2) + delw[0], delw[1])), 4)
        ssx = _scale(_mul(_mul((-om2[0], -om2[1]), _conj(cden)), delw), 1 / _abs2(cden))
        br, sch = 2, zero
    else:
        br, sch, ssx = 3, zero, zero
    # ssx cutoff (|ssx| > 4|I_eps| with wx < 0), compared on squares
    if wx < 0 and _abs2(ssx) > 16 * _abs2(ie):
        ssx = zero
    return br, sch, ssx


def sigma_exact(arrs, nv, occ=1.0):
    """Exact achtemp, asxtemp, acht_n1_loc, achstemp (as Fraction pairs) and,
    for each output, the sum of the magnitudes of its terms (a float scale
    for tolerances)."""
    am, an = arrs["aqsmtemp"], arrs["aqsntemp"]
    ieps, wtil = arrs["I_eps_array"], arrs["wtilde_array"]
    idx, vc, wxs = arrs["inv_igp_index"], arrs["vcoul"], arrs["wx_array"]
    nb, ncouls = am.shape
    ngpown = ieps.shape[0]
    nfreq = len(wxs)
    occ = Fr(occ)
    acht_iw = 2 if nfreq > 2 else nfreq - 1
    zero = (Fr(0), Fr(0))
    ach = [zero] * nfreq
    asx = [zero] * nfreq
    ach_mag = [0.0] * nfreq
    asx_mag = [0.0] * nfreq
    n1loc = [zero] * nb
    n1loc_mag = [0.0] * nb
    achs = zero
    achs_mag = 0.0
    for n1 in range(nb):
        for j in range(ngpown):
            igp = min(int(idx[j]), ncouls - 1)
            vcg = Fr(float(vc[igp]))
            mv1 = _conj(_c(am[n1, igp]))
            # reduce_achstemp (gppComplex.cpp:84-99)
            schs = _c(-ieps[j, igp])
            if _abs2(schs) > TO1 * TO1:
                t = _scale(_mul(_mul(_c(an[n1, igp]), mv1), schs), vcg / 2)
                achs = _add(achs, t)
                achs_mag += _mag(t)
            for iw in range(nfreq):
                wx = Fr(float(wxs[iw]))
                scht = zero
                ssxt = zero
                sch_mag = ssx_mag = 0.0
                for ig in range(ncouls):
                    wt, ie = _c(wtil[j, ig]), _c(ieps[j, ig])
                    a = _mul(mv1, _c(an[n1, ig]))
                    if n1 < nv:
                        _, sch, ssx = pole_terms_exact(wx, wt, ie)
                        scht = _add(scht, _mul(a, sch))
                        ssxt = _add(ssxt, _mul(a, ssx))
                        sch_mag += _mag(_mul(a, sch))
                        ssx_mag += _mag(_mul(a, ssx))
                    else:
                        wdiff = (wx - wt[0], -wt[1])
                        delw = _scale(_mul(wt, _conj(wdiff)), 1 / _abs2(wdiff))
                        t = _mul(a, _mul(delw, ie))
                        scht = _add(scht, t)
                        sch_mag += _mag(t)
                ta = _scale(scht, vcg / 2)
                ach[iw] = _add(ach[iw], ta)
                ach_mag[iw] += sch_mag * float(vcg) / 2
                if iw == acht_iw:
                    n1loc[n1] = _add(n1loc[n1], ta)
                    n1loc_mag[n1] += sch_mag * float(vcg) / 2
                if n1 < nv:
                    tx = _scale(ssxt, occ * vcg)
                    asx[iw] = _add(asx[iw], tx)
                    asx_mag[iw] += ssx_mag * float(occ * vcg)
    return {"achtemp": ach, "asxtemp": asx, "acht_n1_loc": n1loc, "achstemp": achs,
            "achtemp_mag": ach_mag, "asxtemp_mag": asx_mag, "achstemp_mag": achs_mag,
            "acht_n1_loc_mag": n1loc_mag}


def to_complex(p):
    return complex(float(p[0]), float(p[1]))
