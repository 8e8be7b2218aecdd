/*
 * _disptools -- CPython extension with the method table of the reference
 * extension (python_c_extension/_disptools.c:532-540), backed by the HIP
 * library (include/disptools_hip.h).
 *
 * Differences from the reference, all on error paths or defects:
 *   - arrays are checked (dtype, contiguity, shape) and a ValueError is raised
 *     instead of reading foreign memory (reference :342-344 takes raw pointers);
 *   - `strict` is parsed into an int ("p" into a C bool is undefined, :285,309);
 *   - the GIL is released while the device works;
 *   - None is returned with a reference (reference :166,273,421 does not INCREF);
 *   - gpu_id < 0 (the reference's CPU branch, :374-384) calls the same
 *     CPU-named entry points, generate_displacement_gradient / _greedy, which
 *     in this build run on the current HIP device (there is no CPU solver);
 *   - `matching` runs on the HIP device for any gpu_id (the reference rejects
 *     gpu_id >= 0 for it, :368-373).
 * The `shape_descriptor` method (cubeness/octahedroness) is not part of the
 * displacement-generation path and is not provided.
 */
#define PY_SSIZE_T_CLEAN
#define NPY_NO_DEPRECATED_API NPY_1_14_API_VERSION

#include <Python.h>
#include <numpy/arrayobject.h>
#include <string.h>

#include "disptools_hip.h"

#define ALGORITHM_GRADIENT "gradient"
#define ALGORITHM_GREEDY "greedy"
#define ALGORITHM_MATCHING "matching"

static int read_triplet(PyObject *tuple, float *out, const char *name)
{
    if (!PyTuple_Check(tuple)) {
        PyErr_Format(PyExc_ValueError, "%s must be a tuple of floats.", name);
        return -1;
    }
    Py_ssize_t len = PyTuple_Size(tuple);
    if (len != 3) {
        PyErr_Format(PyExc_ValueError, "Wrong number of components (%d) for %s.", (int)len, name);
        return -1;
    }
    for (int i = 0; i < 3; ++i) {
        out[i] = (float)PyFloat_AsDouble(PyTuple_GetItem(tuple, i));
        if (PyErr_Occurred()) {
            PyErr_Clear();
            PyErr_Format(PyExc_ValueError, "%s must be a tuple of floats.", name);
            return -1;
        }
    }
    return 0;
}

/* A C-contiguous, aligned, writeable (if asked) array of the given type and rank. */
static int check_array(PyObject *o, int type, int ndim, int writeable, const char *name)
{
    if (!PyArray_Check(o)) {
        PyErr_Format(PyExc_ValueError, "%s must be a numpy.ndarray.", name);
        return -1;
    }
    PyArrayObject *a = (PyArrayObject *)o;
    if (PyArray_TYPE(a) != type) {
        PyErr_Format(PyExc_ValueError, "%s has the wrong dtype.", name);
        return -1;
    }
    if (PyArray_NDIM(a) != ndim) {
        PyErr_Format(PyExc_ValueError, "%s must have %d dimensions.", name, ndim);
        return -1;
    }
    if (!PyArray_IS_C_CONTIGUOUS(a) || !PyArray_ISALIGNED(a)) {
        PyErr_Format(PyExc_ValueError, "%s must be C-contiguous and aligned.", name);
        return -1;
    }
    if (writeable && !PyArray_ISWRITEABLE(a)) {
        PyErr_Format(PyExc_ValueError, "%s must be writeable.", name);
        return -1;
    }
    return 0;
}

static PyObject *raise_library_error(void)
{
    disptools_error_state *err = disptools_get_error();
    PyErr_Format(PyExc_RuntimeError, "%s (%s:%d): %s\n%s", err->function, err->file, err->line,
                 err->message, err->trace);
    return NULL;
}

static PyObject *method_get_float_type_size(PyObject *self, PyObject *args)
{
    (void)self;
    (void)args;
    return Py_BuildValue("i", get_float_type_size());
}

static PyObject *method_regularise(PyObject *self, PyObject *args)
{
    PyObject *J = NULL;
    float epsilon;
    (void)self;
    if (!PyArg_ParseTuple(args, "Of", &J, &epsilon)) return NULL;
    if (check_array(J, NPY_FLOAT, 3, 1, "Jacobian")) return NULL;
    PyArrayObject *a = (PyArrayObject *)J;
    const size_t nz = PyArray_DIM(a, 0), ny = PyArray_DIM(a, 1), nx = PyArray_DIM(a, 2);
    float *data = (float *)PyArray_DATA(a);
    Py_BEGIN_ALLOW_THREADS
    regularise(nx, ny, nz, data, epsilon);
    Py_END_ALLOW_THREADS
    if (disptools_has_error()) return raise_library_error();
    Py_RETURN_NONE;
}

static PyObject *method_jacobian(PyObject *self, PyObject *args)
{
    float spacing[3];
    PyObject *spacing_tuple = NULL, *field = NULL, *jac = NULL;
    (void)self;
    if (!PyArg_ParseTuple(args, "OOO", &spacing_tuple, &field, &jac)) return NULL;
    if (read_triplet(spacing_tuple, spacing, "spacing")) return NULL;
    if (check_array(field, NPY_FLOAT, 4, 0, "field") || check_array(jac, NPY_FLOAT, 3, 1, "jacobian"))
        return NULL;
    PyArrayObject *f = (PyArrayObject *)field, *J = (PyArrayObject *)jac;
    /* the ITK index order is zyx (reference :261-264) */
    const size_t nz = PyArray_DIM(J, 0), ny = PyArray_DIM(J, 1), nx = PyArray_DIM(J, 2);
    if (PyArray_DIM(f, 0) != 3 || (size_t)PyArray_DIM(f, 1) != nz || (size_t)PyArray_DIM(f, 2) != ny ||
        (size_t)PyArray_DIM(f, 3) != nx) {
        PyErr_SetString(PyExc_ValueError, "field must have shape (3, nz, ny, nx) matching the Jacobian.");
        return NULL;
    }
    const float *fd = (const float *)PyArray_DATA(f);
    float *Jd = (float *)PyArray_DATA(J);
    Py_BEGIN_ALLOW_THREADS
    jacobian_dynamic(nx, ny, nz, spacing[0], spacing[1], spacing[2], fd, Jd);
    Py_END_ALLOW_THREADS
    if (disptools_has_error()) return raise_library_error();
    Py_RETURN_NONE;
}

typedef void (*displacement_fn)(const size_t, const size_t, const size_t, const float, const float,
                                const float, const float *, const bool *, const float, const float, float,
                                const float, const float, const float, const float, const float, const float,
                                const float, const float, const bool, const size_t, float *);

static PyObject *method_displacement(PyObject *self, PyObject *args)
{
    float spacing[3];
    PyObject *spacing_tuple = NULL, *jac = NULL, *mask = NULL, *field = NULL;
    float epsilon, tolerance, eta, eta_max, alpha, beta, gamma, delta, zeta, theta, iota;
    int strict;
    long it_max;
    const char *algorithm;
    int gpu_id;
    (void)self;

    if (!PyArg_ParseTuple(args, "OOOfffffffffffplOsi", &spacing_tuple, &jac, &mask, &epsilon, &tolerance,
                          &eta, &eta_max, &alpha, &beta, &gamma, &delta, &zeta, &theta, &iota, &strict,
                          &it_max, &field, &algorithm, &gpu_id))
        return NULL;

    if (strcmp(algorithm, ALGORITHM_GRADIENT) && strcmp(algorithm, ALGORITHM_GREEDY) &&
        strcmp(algorithm, ALGORITHM_MATCHING)) {
        PyErr_Format(PyExc_ValueError, "Invalid algorithm '%s'.", algorithm);
        return NULL;
    }
    if (read_triplet(spacing_tuple, spacing, "spacing")) return NULL;
    if (check_array(jac, NPY_FLOAT, 3, 0, "Jacobian") || check_array(mask, NPY_BOOL, 3, 0, "mask") ||
        check_array(field, NPY_FLOAT, 4, 1, "field"))
        return NULL;
    PyArrayObject *J = (PyArrayObject *)jac, *M = (PyArrayObject *)mask, *F = (PyArrayObject *)field;
    const size_t nz = PyArray_DIM(J, 0), ny = PyArray_DIM(J, 1), nx = PyArray_DIM(J, 2);
    if ((size_t)PyArray_DIM(M, 0) != nz || (size_t)PyArray_DIM(M, 1) != ny || (size_t)PyArray_DIM(M, 2) != nx) {
        PyErr_SetString(PyExc_ValueError, "The shapes of the Jacobian and the mask must agree.");
        return NULL;
    }
    if (PyArray_DIM(F, 0) != 3 || (size_t)PyArray_DIM(F, 1) != nz || (size_t)PyArray_DIM(F, 2) != ny ||
        (size_t)PyArray_DIM(F, 3) != nx) {
        PyErr_SetString(PyExc_ValueError, "field must have shape (3, nz, ny, nx) matching the Jacobian.");
        return NULL;
    }
    if (it_max < 0) {
        PyErr_SetString(PyExc_ValueError, "it_max must be non-negative.");
        return NULL;
    }

    /* gpu_id >= 0 selects that HIP device and the *_cuda entry points; a
     * negative gpu_id (the reference's default, its CPU branch at
     * _disptools.c:374-384) calls the CPU-named entry points, which run on the
     * current HIP device: this build has no CPU solver. */
    if (gpu_id >= 0 && !set_device(gpu_id)) {
        PyErr_Format(PyExc_ValueError, "Cannot set HIP device %d", gpu_id);
        return NULL;
    }
    displacement_fn fn = NULL;
    if (!strcmp(algorithm, ALGORITHM_GRADIENT)) {
        fn = gpu_id >= 0 ? &generate_displacement_gradient_cuda : &generate_displacement_gradient;
    } else if (!strcmp(algorithm, ALGORITHM_GREEDY)) {
        fn = gpu_id >= 0 ? &generate_displacement_greedy_cuda : &generate_displacement_greedy;
    } else {
        /* CPU-only in the reference (_disptools.c:368-373); here on the GPU */
        fn = &volume_matching_3d;
    }

    const float *Jd = (const float *)PyArray_DATA(J);
    const bool *Md = (const bool *)PyArray_DATA(M);
    float *Fd = (float *)PyArray_DATA(F);
    Py_BEGIN_ALLOW_THREADS
    fn(nx, ny, nz, spacing[0], spacing[1], spacing[2], Jd, Md, epsilon, tolerance, eta, eta_max, alpha, beta,
       gamma, delta, zeta, theta, iota, strict != 0, (size_t)it_max, Fd);
    Py_END_ALLOW_THREADS
    if (disptools_has_error()) return raise_library_error();
    Py_RETURN_NONE;
}

struct module_state {
    PyObject *error;
};

#define GETSTATE(m) ((struct module_state *)PyModule_GetState(m))

static PyObject *error_out(PyObject *self, PyObject *args)
{
    (void)args;
    struct module_state *st = GETSTATE(self);
    PyErr_SetString(st->error, "something bad happened");
    return NULL;
}

static int ext_traverse(PyObject *m, visitproc visit, void *arg)
{
    Py_VISIT(GETSTATE(m)->error);
    return 0;
}

static int ext_clear(PyObject *m)
{
    Py_CLEAR(GETSTATE(m)->error);
    return 0;
}

static PyMethodDef methods[] = {
    {"error_out", (PyCFunction)error_out, METH_NOARGS, NULL},
    {"get_float_type_size", (PyCFunction)method_get_float_type_size, METH_NOARGS,
     "Get the size (in bits) of the floating point type.\n\nReturns: int"},
    {"regularise", (PyCFunction)method_regularise, METH_VARARGS,
     "Apply a lower threshold to the Jacobian (in place).\n\n"
     "Parameters\n    Jacobian (float32 ndarray, zyx)\n    epsilon (float)\n"},
    {"jacobian", (PyCFunction)method_jacobian, METH_VARARGS,
     "Compute the Jacobian of a displacement field on the current HIP device.\n\n"
     "Parameters\n    spacing (tuple of three floats)\n"
     "    field (float32 ndarray, dzyx)\n    result Jacobian (float32 ndarray, zyx, written)\n"},
    {"displacement", (PyCFunction)method_displacement, METH_VARARGS,
     "Compute a displacement with given Jacobian (field modified in place).\n\n"
     "Parameters\n    spacing, Jacobian, mask, epsilon, tolerance, eta, eta_max, alpha, beta,\n"
     "    gamma, delta, zeta, theta, iota, strict, it_max, field, algorithm, gpu_id\n"},
    {NULL, NULL, 0, NULL}};

static struct PyModuleDef moduledef = {
    PyModuleDef_HEAD_INIT, "_disptools",
    "Python C extension module.\n\nThis module wraps the HIP library for the generation of "
    "displacement fields.",
    sizeof(struct module_state), methods, NULL, ext_traverse, ext_clear, NULL};

PyMODINIT_FUNC PyInit__disptools(void)
{
    PyObject *module = PyModule_Create(&moduledef);
    if (module == NULL) return NULL;
    struct module_state *st = GETSTATE(module);
    st->error = PyErr_NewException("_disptools.Error", NULL, NULL);
    if (st->error == NULL) {
        Py_DECREF(module);
        return NULL;
    }
    import_array();
    return module;
}
