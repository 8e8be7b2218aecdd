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
  reference also passes through ``drawing.scale_image``) is built and solved
  on the device by ``dt_displacement_levels``: the recursive Gaussian of
  sigma 2 (physical units) that SimpleITK's SmoothingRecursiveGaussian
  documents (Deriche's fourth-order filter, restated from the published
  method: SimpleITK is not installed here, so agreement with it is pinned by
  the filter's mathematics, not bit for bit), then the resampling of
  ResampleImageFilter (same origin and extent, linear or nearest-neighbour,
  0 outside).  All levels, masks and intermediate fields stay in device
  memory; numpy arrays in and out, no torch.
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


def _pyramid_displacement(J, M, guess, spacing, levels, it_max, parameters):
    """Multi-resolution solve (reference displacements.py:449-483) on the
    device through dt_displacement_levels: every coarser level is scaled from
    the finest padded target and mask; each level starts from the previous
    level's field scaled onto its grid (the initial guess, if any, for the
    coarsest).  Host arrays in and out; the library allocates and frees the
    device buffers itself."""
    from . import native
    gpu_id = int(parameters["gpu_id"])
    if gpu_id >= 0 and not native.set_device(gpu_id):
        raise ValueError("HIP device %d is not available" % gpu_id)
    # gpu_id < 0: the current device (the reference's CPU solver, which this
    # build does not have, runs there instead)
    for size in pyramid_sizes(tuple(reversed(J.shape)), levels):
        logging.info('Size %s' % str(size))
    algorithm = parameters["algorithm"]
    params = {k: v for k, v in parameters.items() if k not in ("algorithm", "gpu_id")}
    if algorithm == "matching":
        params = {k: params[k] for k in ("epsilon", "delta", "zeta")}
    # the reference casts the mask to float before scaling it (displacements.py:407)
    field, _ = native.displacement_levels(algorithm, J, np.asarray(M, np.float32), guess, spacing, levels,
                                          list(it_max), **params)
    return field


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
