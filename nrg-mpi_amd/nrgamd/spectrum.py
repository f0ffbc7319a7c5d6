"""Coarse log-spaced view A(ω) of the binned CFS spectrum.

The spectral weights of `spikeCFS` / `spikeZeroTemperature` are binned on
the reference's native mesh: point i (0-based) sits at
ω_i = ±10^(log10Emin + i/meshOfBin), log10Emin = −15, meshOfBin = 1000
(`module.F90:964-970`, `main.F90:142-156`), and `binningSpike`
(`module.F90:1210-1254`) splits every spike's weight between the two mesh
points around it.  A native bin therefore holds a weight, not a density.

`log_view` is the coarse spectral function used for reporting an
Anderson-model run on a few hundred ω points (the "200 log-spaced bins"
view): native point i stands for the log-cell
[log10Emin + (i − ½)/mesh, log10Emin + (i + ½)/mesh), and the coarse bins
are unions of whole cells, `nbins` of them, evenly spaced in log10 ω over
[lo, hi) (default −8 .. 0: 25 bins per decade, 40 native cells each).  Each
coarse bin's weight is the sum of its native weights, its A(ω) that weight
over the bin's width in ω, reported at the geometric centre of the bin.  The
view is an exact regrouping: no interpolation and no smoothing, so the
coarse weights plus the weight outside [lo, hi) add up to the native total.
"""
import numpy as np

LOG10_EMIN = -15
MESH_OF_BIN = 1000


def _cells(lo, hi, nbins, mesh):
    span = (hi - lo) * mesh
    if span <= 0 or round(span) != span or int(round(span)) % nbins:
        raise ValueError(f"[{lo}, {hi}) does not split into {nbins} bins of whole native cells at {mesh} per decade")
    width = int(round(span)) // nbins
    first = int(round((lo - LOG10_EMIN) * mesh))
    if first < 0:
        raise ValueError(f"lower edge 10^{lo} is below the native mesh (10^{LOG10_EMIN})")
    return first, width


def log_view(bins, lo=-8, hi=0, nbins=200, mesh=MESH_OF_BIN):
    """Coarse view of native bins.

    bins: (nmat, nbin) or (nbin,) native weights of one sign (the second
    column of `binnedSpikePositive` / `binnedSpikeNegative` per matrix).
    Returns (omega, A, weight, outside): the bin centres |ω| (nbins), A(ω)
    (nmat, nbins), the coarse weights (nmat, nbins), and the native weight
    outside [10^lo, 10^hi) per matrix (nmat,).
    """
    b = np.atleast_2d(np.asarray(bins, dtype=np.float64))
    first, width = _cells(lo, hi, nbins, mesh)
    last = first + width * nbins
    if last > b.shape[1]:
        raise ValueError(f"upper edge 10^{hi} is above the native mesh of {b.shape[1]} points")
    weight = b[:, first:last].reshape(b.shape[0], nbins, width).sum(axis=2)
    k = np.arange(nbins + 1)
    edges = 10.0 ** (LOG10_EMIN + (first + k * width - 0.5) / mesh)
    omega = np.sqrt(edges[:-1] * edges[1:])
    A = weight / np.diff(edges)[None, :]
    outside = b.sum(axis=1) - weight.sum(axis=1)
    return omega, A, weight, outside
