"""Seeded random-ellipse phantoms, rasterised on the GPU.

Phantom i of a seeded set draws its parameters from
``numpy.random.default_rng([seed, i])``:

* ellipse count K ~ U{10, ..., 20}
* centre sqrt(U) * (cos 2 pi U', sin 2 pi U')  (uniform in the unit disk)
* semi-axes U[0.05, 0.5] each, rotation U[0, pi), intensity U[-0.5, 1.0]

Pixel centres sit at -1 + (2j + 1) / N along x and 1 - (2i + 1) / N along y
(row 0 at the top).  Each pixel sums the intensities of the ellipses whose
closed interior contains its centre, evaluated in float64 and cast to the
requested dtype at the end.  The parameters come from NumPy on the host (a
few hundred numbers); the N x N rasterisation runs on the device.
"""

import numpy as np

from . import _device as dev


def ellipse_params(seed, index):
    rng = np.random.default_rng([seed, index])
    k = int(rng.integers(10, 21))
    r = np.sqrt(rng.random(k))
    t = 2 * np.pi * rng.random(k)
    cx, cy = r * np.cos(t), r * np.sin(t)
    ax = rng.uniform(0.05, 0.5, k)
    ay = rng.uniform(0.05, 0.5, k)
    rot = rng.uniform(0.0, np.pi, k)
    val = rng.uniform(-0.5, 1.0, k)
    return np.stack([cx, cy, ax, ay, rot, val], axis=1)


def phantoms(count, n, *, seed=0, dtype="float32", device=None, start=0):
    """``count`` phantoms of side ``n`` as a (count, n, n) device tensor."""
    torch = dev.torch()
    if device is None:
        dev.require_gpu()
        device = torch.device("cuda", torch.cuda.current_device())
    coords = -1.0 + (2.0 * torch.arange(n, device=device, dtype=torch.float64) + 1.0) / n
    xs = coords.view(1, n)
    ys = -coords.view(n, 1)
    out = torch.empty((count, n, n), device=device, dtype=getattr(torch, dtype))
    for i in range(count):
        p = torch.as_tensor(ellipse_params(seed, start + i), device=device)
        img = torch.zeros((n, n), device=device, dtype=torch.float64)
        for cx, cy, ax, ay, rot, val in p:
            c, s = torch.cos(rot), torch.sin(rot)
            dx, dy = xs - cx, ys - cy
            u = (dx * c + dy * s) / ax
            v = (-dx * s + dy * c) / ay
            img += torch.where(u * u + v * v <= 1.0, val, torch.zeros((), device=device,
                                                                       dtype=torch.float64))
        out[i] = img.to(out.dtype)
    return out


def phantoms_numpy(count, n, *, seed=0, dtype="float32", start=0):
    """Host copy of :func:`phantoms` (same values)."""
    return phantoms(count, n, seed=seed, dtype=dtype, start=start).cpu().numpy()
