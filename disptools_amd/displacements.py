"""Drop-in counterparts of ``disptools.displacements.displacement`` and ``jacobian``.

Reference: ``disptools/displacements.py:55-92`` (jacobian), ``:96-202``
(_displacement), ``:238-485`` (displacement).  Arguments keep their names,
defaults and meaning.  Images are ``SimpleITK.Image`` objects when SimpleITK is
installed, or numpy arrays: a scalar volume ``[z, y, x]`` and a vector field
``[z, y, x, 3]`` (ITK array order), with the voxel spacing ``(dx, dy, dz)``
given by the keyword ``spacing`` (default 1).

Behaviour of this build that differs from the reference:

* the solve runs on a HIP device; ``gpu_id`` must be >= 0 (``-1`` selects the
  reference's CPU solver, which this build does not have, and raises
  ``ValueError`` from the extension);
* ``levels > 1`` needs the reference's SimpleITK pyramid
  (``drawing.scale_image``: recursive Gaussian + linear resampling), which is
  not reproduced: it raises ``NotImplementedError``;
* an ``initial_guess`` is padded by edge replication to the padded grid; the
  reference resamples it onto that grid with ``scale_image`` (smoothing it).
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

    if levels != 1:
        raise NotImplementedError("levels > 1 needs the SimpleITK pyramid of the reference "
                                  "(drawing.scale_image), which this build does not reproduce")
    if isinstance(it_max, list):
        if len(it_max) != levels:
            raise ValueError('len(it_max) should equal the value of `levels`')
        it_max = it_max[0]

    # one voxel of zero-flux padding at least (reference :443-447)
    pad = pad + 1 if pad > 0 else 1
    if redistribute:
        J = redistribute_volume_change(J, M).astype(np_float_type)
    M = np.pad(M, pad, mode="constant", constant_values=0)
    J = np.pad(J, pad, mode="edge")
    if guess is not None:
        guess = np.pad(guess, [(pad, pad)] * 3 + [(0, 0)], mode="edge")

    logging.info('Size %s' % str(tuple(reversed(J.shape))))
    field = _displacement(J, M != 0, guess, sp, it_max=it_max, **parameters)
    field = field[pad:-pad, pad:-pad, pad:-pad, :]
    if sitk_in:
        out = sitk.GetImageFromArray(np.ascontiguousarray(field), isVector=True)
        out.CopyInformation(jacobian)
        return out
    return np.ascontiguousarray(field)
