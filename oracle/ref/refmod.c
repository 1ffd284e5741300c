/* Test-only harness: exposes the reference's own C objects (Operator,
 * Transformer, Spline, LFC, spline_to_grid) as a Python module named
 * _gpawref, so the oracle and the golden-fixture script can be checked
 * against the reference's compiled C code in this container.
 *
 * Nothing here computes anything: every entry point is the reference's
 * constructor, compiled from the reference sources where they lie.  This
 * file only replaces the reference's module table (c/_gpaw.c:125-237,
 * PyInit__gpaw c/_gpaw.c:328), which cannot be compiled here because it
 * also registers the XC functionals (libxc is not installed).
 */
#include <Python.h>
#define PY_ARRAY_UNIQUE_SYMBOL GPAW_ARRAY_API
#include <numpy/arrayobject.h>

PyObject* NewOperatorObject(PyObject* self, PyObject* args);
PyObject* NewTransformerObject(PyObject* self, PyObject* args);
PyObject* NewSplineObject(PyObject* self, PyObject* args);
PyObject* NewLFCObject(PyObject* self, PyObject* args);
PyObject* spline_to_grid(PyObject* self, PyObject* args);

extern PyTypeObject OperatorType;
extern PyTypeObject TransformerType;
extern PyTypeObject SplineType;
extern PyTypeObject LFCType;

static PyMethodDef methods[] = {
    {"Operator", NewOperatorObject, METH_VARARGS, 0},
    {"Transformer", NewTransformerObject, METH_VARARGS, 0},
    {"Spline", NewSplineObject, METH_VARARGS, 0},
    {"LFC", NewLFCObject, METH_VARARGS, 0},
    {"spline_to_grid", spline_to_grid, METH_VARARGS, 0},
    {0, 0, 0, 0}};

static struct PyModuleDef moduledef = {
    PyModuleDef_HEAD_INIT, "_gpawref", "reference C subset (test only)", -1,
    methods, 0, 0, 0, 0};

PyMODINIT_FUNC PyInit__gpawref(void)
{
    if (PyType_Ready(&OperatorType) < 0) return 0;
    if (PyType_Ready(&TransformerType) < 0) return 0;
    if (PyType_Ready(&SplineType) < 0) return 0;
    if (PyType_Ready(&LFCType) < 0) return 0;
    import_array1(0);
    return PyModule_Create(&moduledef);
}
