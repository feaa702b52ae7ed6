"""Statistics that compare two erosion runs which cannot agree cell for cell
(different particle order and random streams): the height distribution,
the dominant dune wavelength and the total mass.

These implement the comparison the batched method is held to against the
reference's sequential run (`World::erode`, world.hpp:36-77, one particle at
a time).  Pure numpy on height maps.
"""
import numpy as np


def wasserstein1(a, b):
    """Earth mover's distance between the value distributions of two maps of
    equal size (mean absolute difference of the sorted values)."""
    a = np.sort(np.asarray(a, np.float64).ravel())
    b = np.sort(np.asarray(b, np.float64).ravel())
    if a.size != b.size:
        q = np.linspace(0.0, 1.0, min(a.size, b.size))
        a, b = np.quantile(a, q), np.quantile(b, q)
    return float(np.mean(np.abs(a - b)))


def height_histogram(h, bins=64, lo=None, hi=None):
    h = np.asarray(h, np.float64).ravel()
    lo = float(h.min()) if lo is None else lo
    hi = float(h.max()) if hi is None else hi
    counts, edges = np.histogram(h, bins=bins, range=(lo, hi))
    return counts / max(1, h.size), edges


def radial_spectrum(h):
    """Radially averaged power spectrum of a (periodic) height map; index k is
    the number of periods across the map."""
    h = np.asarray(h, np.float64)
    f = np.fft.fft2(h - h.mean())
    p = np.abs(f) ** 2
    nx, ny = h.shape
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    k = np.sqrt(kx[:, None] ** 2 + ky[None, :] ** 2)
    kb = np.rint(k).astype(np.int64)
    kmax = min(nx, ny) // 2
    sel = (kb >= 1) & (kb <= kmax)
    spec = np.bincount(kb[sel], weights=p[sel], minlength=kmax + 1)
    cnt = np.bincount(kb[sel], minlength=kmax + 1)
    with np.errstate(invalid="ignore", divide="ignore"):
        spec = np.where(cnt > 0, spec / np.maximum(cnt, 1), 0.0)
    return spec


def dune_wavelength(h, kmin=2):
    """Wavelength (cells) at the peak of the radially averaged power spectrum
    of the erosion change, ignoring wavenumbers below `kmin`."""
    spec = radial_spectrum(h)
    k = int(np.argmax(spec[kmin:]) + kmin)
    return min(np.asarray(h).shape) / k, k


def relative_mass_difference(a, b):
    ma = float(np.asarray(a, np.float64).sum())
    mb = float(np.asarray(b, np.float64).sum())
    return abs(ma - mb) / max(abs(ma), abs(mb), 1e-300)


def compare(h0, seq, batched):
    """Statistics of two runs from the same initial map h0."""
    ds, db = np.asarray(seq, np.float64) - h0, np.asarray(batched, np.float64) - h0
    return {
        "w1_height": wasserstein1(seq, batched),
        "w1_change": wasserstein1(ds, db),
        "rms_change_seq": float(np.sqrt(np.mean(ds ** 2))),
        "rms_change_batched": float(np.sqrt(np.mean(db ** 2))),
        "wavelength_seq": dune_wavelength(ds)[0],
        "wavelength_batched": dune_wavelength(db)[0],
        "mass_rel_diff": relative_mass_difference(seq, batched),
        "mass_rel_loss_seq": relative_mass_difference(h0, seq),
        "mass_rel_loss_batched": relative_mass_difference(h0, batched),
    }
