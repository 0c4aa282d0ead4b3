"""Shared comparison of a GPU fit with the oracle (
coordinates within 1e-4 px (spline units for z, w, u; absolute for g2),
photons / background within 1e-4 relative, CRLB within 1e-4 relative,
iteration counts equal)."""
import numpy as np

COORD = {6: [0, 1, 4], 7: [0, 1, 4, 5], 8: [0, 1, 4, 5, 6, 7]}
REL = {6: [2, 3], 7: [2, 3], 8: [2, 3]}


def compare(ft, P, CR, LL, Po, CRo, LLo, atol=1e-4, rtol=1e-4):
    NV = CR.shape[1]
    P = np.asarray(P, np.float64)
    Po = np.asarray(Po, np.float64)
    dcoord = np.abs(P[:, COORD[ft]] - Po[:, COORD[ft]]).max(1)
    drel = (np.abs(P[:, REL[ft]] - Po[:, REL[ft]]) / np.maximum(np.abs(Po[:, REL[ft]]), 1e-12)).max(1)
    dcr = (np.abs(CR - CRo) / np.maximum(np.abs(CRo), 1e-30)).max(1)
    same_kk = P[:, NV] == Po[:, NV]
    ok = (dcoord <= atol) & (drel <= rtol) & (dcr <= rtol) & same_kk
    return dict(n=len(P), agree=float(ok.mean()), same_kk=float(same_kk.mean()),
                p_ok=float(((dcoord <= atol) & (drel <= rtol)).mean()),
                crlb_ok=float((dcr <= rtol).mean()),
                p99_dcoord=float(np.percentile(dcoord, 99)), max_dcoord=float(dcoord.max()),
                p99_drel=float(np.percentile(drel, 99)), p99_dcrlb=float(np.percentile(dcr, 99)),
                finite=bool(np.all(np.isfinite(P)) and np.all(np.isfinite(CR))))
