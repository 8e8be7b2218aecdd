"""Drop-in counterparts of ``disptools.displacements.displacement`` and ``jacobian``.

Reference: ``disptools/displacements.py:55-92`` (jacobian), ``:96-202``
(_displacement), ``:238-485`` (displacement).  Arguments keep their names,
defaults and meaning.  Images are ``SimpleITK.Image`` objects when SimpleITK is
installed, or numpy arrays: a scalar volume ``[z, y, x]`` and a vector field
``[z, y, x, 3]`` (ITK array order), with the voxel spacing ``(dx, dy, dz)``
given by the keyword ``spacing`` (default 1).

Behaviour of this build that differs from the reference:

* the solve always runs on a HIP device: ``gpu_id >= 0`` selects the device,
  the default ``gpu_id=-1`` (the reference's CPU solver) runs on the current
  HIP device through the reference's CPU-named entry points
  (``generate_displacement_gradient`` / ``_greedy``); there is no CPU solver;
* the image pyramid (``levels > 1``, and an ``initial_guess``, which the
  reference also passes through ``drawing.scale_image``) is built on the
  device by ``dt_scale_volume_device``: a Gaussian of sigma 2 (physical
  units) with taps out to 4 sigma and replicated edges, then the same
  resampling as SimpleITK's ResampleImageFilter (same origin and extent,
  linear or nearest-neighbour, 0 outside).  The reference smooths with
  SimpleITK's recursive (IIR) Gaussian, which is not installed here; the
  levels therefore agree with the reference's only to that filter's
  approximation of a Gaussian (parity unpinned).  All levels, masks and
  intermediate fields stay in device memory.
"""

import logging
from typing import List, Union

import numpy as np

from . import load_extension

try:  # optional, as the image container of the reference API
    import SimpleITK as sitk  # noqa: F401
    HAVE_SITK = True
except ImportError:  # pragma: no cover - SimpleITK is not installed in this image
    sitk = None
    HAVE_SITK = False

_ext = load_extension()
np_float_type = np.float32 if _ext.get_float_type_size() == 32 else np.float64


def _is_sitk(img):
    return HAVE_SITK and isinstance(img, sitk.Image)


def _scalar_array(img, spacing):
    if _is_sitk(img):
        return sitk.GetArrayFromImage(img), tuple(img.GetSpacing())
    a = np.asarray(img)
    if a.ndim != 3:
        raise Exception("Wrong jacobian dimensionality")
    return a, tuple(float(s) for s in spacing)


def jacobian(field, *, spacing=(1.0, 1.0, 1.0)):
    r"""Compute the Jacobian map of a vector field (reference displacements.py:55-92).

    Parameters
    ----------
    field : sitk.Image or numpy.ndarray [z, y, x, 3]
        Input vector field.
    spacing : tuple
        Voxel spacing (dx, dy, dz) of a numpy input (a SimpleITK image carries its own).

    Returns
    -------
    The Jacobian map, of the same kind as the input (numpy [z, y, x] for numpy input).
    """
    if _is_sitk(field):
        image = sitk.GetArrayViewFromImage(sitk.Cast(field, sitk.sitkVectorFloat32))
        sp = tuple(field.GetSpacing())
    else:
        image = np.asarray(field)
        sp = tuple(float(s) for s in spacing)
    if 4 != len(image.shape) or 3 != image.shape[3]:
        raise Exception("Wrong input dimensionality")
    shape = image.shape[0:3]
    data = np.empty((3, *shape), dtype=np_float_type)
    for c in range(3):
        data[c] = image[..., c]
    result = np.zeros(shape, dtype=np_float_type)
    _ext.jacobian(sp, data, result)
    if _is_sitk(field):
        out = sitk.GetImageFromArray(result)
        out.CopyInformation(field)
        return out
    return result


def regularise(jacobian, epsilon: float = 1e-5):
    r"""Replace Jacobian values below ``epsilon`` by ``epsilon`` (reference displacements.py:22-52)."""
    if _is_sitk(jacobian):
        result = sitk.Cast(jacobian, sitk.sitkFloat32)
        a = sitk.GetArrayFromImage(result)
        _ext.regularise(a, float(epsilon))
        out = sitk.GetImageFromArray(a)
        out.CopyInformation(jacobian)
        return out
    a = np.array(jacobian, dtype=np_float_type, order="C", copy=True)
    if a.ndim != 3:
        raise Exception("Wrong jacobian dimensionality")
    _ext.regularise(a, float(epsilon))
    return a


def redistribute_volume_change(image, mask):
    r"""Move the volume change of the body onto the background so the total change
    is zero (reference displacements.py:205-235)."""
    data = np.array(image, dtype=np.float64 if np.asarray(image).dtype == np.float64 else np_float_type)
    index = np.asarray(mask)
    correction = -(np.sum(data) - data.size) / (data.size - np.count_nonzero(index))
    data[index == 0.0] += correction
    return data


def _displacement(jacobian, mask, initial_guess, spacing, *, epsilon, tolerance, it_max, alpha, beta,
                  gamma, delta, zeta, theta, iota, strict, eta, eta_max, algorithm, gpu_id):
    """One resolution level (reference displacements.py:96-202)."""
    J = np.array(jacobian, dtype=np_float_type, order="C", copy=True)
    m = np.array(mask, dtype=bool, order="C", copy=True)
    if m.shape != J.shape:
        raise Exception("The shapes of the Jacobian and the mask must agree")
    if initial_guess is not None:
        data = np.asarray(initial_guess)
        if data.shape[0:3] != J.shape:
            raise Exception("The shapes of the Jacobian and the initial guess must agree")
        field = np.empty((3, *J.shape), dtype=np_float_type)
        for c in range(3):
            field[c] = data[..., c]
    else:
        field = np.zeros((3, *J.shape), dtype=np_float_type)
    _ext.displacement(tuple(float(s) for s in spacing), J, m, epsilon, tolerance, eta, eta_max, alpha,
                      beta, gamma, delta, zeta, theta, iota, bool(strict), int(it_max), field, algorithm,
                      int(gpu_id))
    out = np.empty((*J.shape, 3), dtype=np_float_type)
    for c in range(3):
        out[..., c] = field[c]
    return out


def pyramid_sizes(size_xyz, levels):
    """Grid sizes (x, y, z) of the pyramid levels, finest first: each level
    halves the previous one (reference displacements.py:454-458)."""
    sizes = [tuple(int(v) for v in size_xyz)]
    for _ in range(1, levels):
        sizes.append(tuple(v // 2 for v in sizes[-1]))
    return sizes


def _scale(d_in, spacing, size_xyz, nearest=False, sigma=2.0):
    """drawing.scale_image on a device volume [ncomp, nz, ny, nx]: returns
    the resampled volume (same origin and physical extent) and its spacing."""
    import torch
    from . import native
    ncomp, nz, ny, nx = d_in.shape
    mx, my, mz = size_xyz
    if min(mx, my, mz) < 1:
        raise ValueError("pyramid level of size %r is empty" % (size_xyz,))
    out = torch.empty((ncomp, mz, my, mx), dtype=torch.float32, device=d_in.device)
    native.scale_volume_device(nx, ny, nz, ncomp, spacing, d_in.data_ptr(), (mx, my, mz), sigma, nearest,
                               out.data_ptr(), torch.cuda.current_stream(d_in.device).cuda_stream)
    new_spacing = tuple(s * n / m for s, n, m in zip(spacing, (nx, ny, nz), (mx, my, mz)))
    return out, new_spacing


def _pyramid_displacement(J, M, guess, spacing, levels, it_max, parameters):
    """Multi-resolution solve on the device (reference displacements.py:449-483):
    every coarser level is scaled from the finest padded target and mask;
    each level starts from the previous level's field scaled onto its grid
    (the initial guess, if any, for the coarsest)."""
    import torch
    from . import native
    gpu_id = int(parameters["gpu_id"])
    if gpu_id >= 0 and not native.set_device(gpu_id):
        raise ValueError("HIP device %d is not available" % gpu_id)
    # gpu_id < 0: the current device (the reference's CPU solver, which this
    # build does not have, runs there instead)
    dev = torch.device("cuda", gpu_id if gpu_id >= 0 else torch.cuda.current_device())
    nz, ny, nx = J.shape
    sizes = pyramid_sizes((nx, ny, nz), levels)
    dJ = torch.from_numpy(np.ascontiguousarray(J, np.float32)[None]).to(dev)
    dM = torch.from_numpy(np.ascontiguousarray(M != 0, np.float32)[None]).to(dev)
    J_levels, M_levels, sp_levels = [dJ[0]], [dM[0]], [tuple(spacing)]
    for size in sizes[1:]:
        j, sp_l = _scale(dJ, spacing, size)
        m, _ = _scale(dM, spacing, size, nearest=True)
        J_levels.append(j[0])
        M_levels.append(m[0])
        sp_levels.append(sp_l)
    field, f_sp = None, tuple(spacing)
    if guess is not None:
        # [z, y, x, 3] -> [3, z, y, x] on the device
        g = torch.from_numpy(np.ascontiguousarray(guess, np.float32)).to(dev)
        field = g.permute(3, 0, 1, 2).contiguous()
    params = {k: v for k, v in parameters.items() if k not in ("algorithm", "gpu_id")}
    stream = torch.cuda.current_stream(dev).cuda_stream
    # it_max[k] belongs to the k-th level from the coarsest (reference :476 zips
    # the reversed pyramid with it_max)
    for k, level in enumerate(reversed(range(levels))):
        mx, my, mz = sizes[level]
        logging.info('Size %s' % str(sizes[level]))
        if field is not None:
            field, _ = _scale(field, f_sp, sizes[level])
        else:
            field = torch.zeros((3, mz, my, mx), dtype=torch.float32, device=dev)
        T = J_levels[level].contiguous()
        Mb = (M_levels[level] != 0).to(torch.uint8).contiguous()
        field = field.contiguous()
        if parameters["algorithm"] == "matching":
            native.matching_device(mx, my, mz, sp_levels[level], T.data_ptr(), field.data_ptr(), stream=stream,
                                   it_max=it_max[k], **params)
        else:
            native.solve_device(parameters["algorithm"], mx, my, mz, sp_levels[level], T.data_ptr(),
                                Mb.data_ptr(), field.data_ptr(), stream=stream, it_max=it_max[k], **params)
        f_sp = sp_levels[level]
    # [3, z, y, x] -> [z, y, x, 3] on the device, then one download
    return field.permute(1, 2, 3, 0).contiguous().cpu().numpy()


def displacement(
        jacobian,
        *,
        levels: int = 1,
        pad: int = 0,
        redistribute: bool = False,
        mask=None,
        initial_guess=None,
        epsilon: float = 9.99e-4,
        tolerance: float = 0.2,
        it_max: Union[int, List[int]] = 50000,
        alpha: float = 1.2,
        beta: float = 0.5,
        gamma: float = 0.1,
        delta: float = 1e-3,
        zeta: float = 10.0,
        theta: float = 1e-6,
        iota: float = 1e-9,
        strict: bool = False,
        eta: float = 0.1,
        eta_max: float = 0.4,
        algorithm: str = 'gradient',
        gpu_id: int = -1,
        spacing=(1.0, 1.0, 1.0),
        ):
    r"""Generate a displacement field that realises a given Jacobian
    (reference displacements.py:238-485; same parameters and defaults).

    ``spacing`` applies to numpy inputs only.  Returns a vector field of the
    same kind as the input: a SimpleITK image, or a numpy array [z, y, x, 3].
    """
    parameters = dict(epsilon=epsilon, tolerance=tolerance, alpha=alpha, beta=beta, gamma=gamma,
                      delta=delta, zeta=zeta, theta=theta, iota=iota, strict=strict, eta=eta,
                      eta_max=eta_max, algorithm=algorithm, gpu_id=gpu_id)
    sitk_in = _is_sitk(jacobian)
    J, sp = _scalar_array(jacobian, spacing)
    J = J.astype(np_float_type)
    if mask is None:
        M = np.ones(J.shape, dtype=np_float_type)
    else:
        M = sitk.GetArrayFromImage(mask) if _is_sitk(mask) else np.asarray(mask)
        if M.shape != J.shape:
            raise Exception("The shapes of the Jacobian and the mask must agree")
    guess = None
    if initial_guess is not None:
        guess = sitk.GetArrayFromImage(initial_guess) if _is_sitk(initial_guess) else np.asarray(initial_guess)
        if guess.shape[0:3] != J.shape:
            raise Exception("The shapes of the Jacobian and the initial guess must agree")

    if not isinstance(levels, int) or levels < 1:
        raise ValueError("levels must be a positive integer")
    if not isinstance(it_max, list):
        it_max = [it_max for _ in range(levels)]
    elif len(it_max) != levels:
        raise ValueError('len(it_max) should equal the value of `levels`')

    # one voxel of zero-flux padding at least (reference :443-447)
    pad = pad + 1 if pad > 0 else 1
    if redistribute:
        J = redistribute_volume_change(J, M).astype(np_float_type)
    M = np.pad(M, pad, mode="constant", constant_values=0)
    J = np.pad(J, pad, mode="edge")

    if levels == 1 and guess is None:
        logging.info('Size %s' % str(tuple(reversed(J.shape))))
        field = _displacement(J, M != 0, None, sp, it_max=it_max[0], **parameters)
    else:
        field = _pyramid_displacement(J, M, guess, sp, levels, it_max, parameters)
    field = field[pad:-pad, pad:-pad, pad:-pad, :]
    if sitk_in:
        out = sitk.GetImageFromArray(np.ascontiguousarray(field), isVector=True)
        out.CopyInformation(jacobian)
        return out
    return np.ascontiguousarray(field)
