"""Deterministic test phantoms (numpy): rasterised ellipses and Gaussian blobs
with their analytic parallel-beam line integrals."""
import numpy as np

# Modified Shepp-Logan (Toft): centre x, centre y, semi-axis a, semi-axis b,
# rotation (degrees), additive intensity.
SHEPP_LOGAN = [
    (0.0, 0.0, 0.69, 0.92, 0.0, 1.0),
    (0.0, -0.0184, 0.6624, 0.874, 0.0, -0.8),
    (0.22, 0.0, 0.11, 0.31, -18.0, -0.2),
    (-0.22, 0.0, 0.16, 0.41, 18.0, -0.2),
    (0.0, 0.35, 0.21, 0.25, 0.0, 0.1),
    (0.0, 0.1, 0.046, 0.046, 0.0, 0.1),
    (0.0, -0.1, 0.046, 0.046, 0.0, 0.1),
    (-0.08, -0.605, 0.046, 0.023, 0.0, 0.1),
    (0.0, -0.605, 0.023, 0.023, 0.0, 0.1),
    (0.06, -0.605, 0.023, 0.046, 0.0, 0.1),
]


def raster_ellipses(n, ellipses, extent=None, clamp=False):
    """Sum of ellipses on the [-1, 1]^2 square rasterised at the pixel centres
    of an n x n grid (y axis pointing up); extent restricts the phantom to the
    central `extent` pixels of the grid."""
    extent = n if extent is None else extent
    c = (np.arange(n) + 0.5 - n / 2.0) / (extent / 2.0)
    X, Y = np.meshgrid(c, -c)
    img = np.zeros((n, n), np.float64)
    for (x0, y0, a, b, phi, val) in ellipses:
        t = np.deg2rad(phi)
        xr = (X - x0) * np.cos(t) + (Y - y0) * np.sin(t)
        yr = -(X - x0) * np.sin(t) + (Y - y0) * np.cos(t)
        img[(xr / a) ** 2 + (yr / b) ** 2 <= 1.0] += val
    if clamp:
        np.maximum(img, 0, out=img)
    return img.astype(np.float32)


def random_ellipses(seed, count=20):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        r = np.sqrt(rng.random())
        ang = 2 * np.pi * rng.random()
        out.append((r * np.cos(ang), r * np.sin(ang), rng.uniform(0.02, 0.3), rng.uniform(0.02, 0.3),
                    np.rad2deg(rng.uniform(0, np.pi)), rng.uniform(-0.2, 0.5)))
    return out


def ellipse_phantom(n, seed=7, count=20):
    return raster_ellipses(n, random_ellipses(seed, count), clamp=True)


def shepp_logan(n, extent=None):
    return raster_ellipses(n, SHEPP_LOGAN, extent=extent)


def gaussian_blobs(n, blobs):
    """Image of Gaussian blobs; blobs = [(X0, Y0, sigma, amp)] in pixel units
    relative to the image centre (X right along columns, Y down along rows)."""
    X = np.arange(n) + 0.5 - n / 2.0
    XX, YY = np.meshgrid(X, X)
    img = np.zeros((n, n))
    for (x0, y0, s, amp) in blobs:
        img += amp * np.exp(-((XX - x0) ** 2 + (YY - y0) ** 2) / (2 * s * s))
    return img


def gaussian_blobs_radon(blobs, theta, t):
    """Analytic line integral of gaussian_blobs along X cos(theta) - Y sin(theta) = t."""
    th = np.asarray(theta)[:, None]
    tt = np.asarray(t)
    out = np.zeros(np.broadcast(th, tt).shape)
    for (x0, y0, s, amp) in blobs:
        t0 = x0 * np.cos(th) - y0 * np.sin(th)
        out += amp * np.sqrt(2 * np.pi) * s * np.exp(-((tt - t0) ** 2) / (2 * s * s))
    return out
