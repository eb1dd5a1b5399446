// CPython module `libMHCUDA`: the Python face of the engine, with the same five
// functions, argument formats, dtype rules, row slicing and exception mapping
// as the reference's python.cc:
//   minhash_cuda_init(dim, samples, seed=time(), deferred=False, devices=0,
//                     verbosity=0)                       python.cc:98-143
//   minhash_cuda_retrieve_vars(gen)                      python.cc:145-195
//   minhash_cuda_assign_vars(gen, rs, ln_cs, betas)      python.cc:197-272
//   minhash_cuda_calc(gen, matrix, row_start=0,
//                     row_finish=0xffffffff)             python.cc:274-399
//     (matrices of 2^32 or more nonzeros, which the reference's uint32
//     force-cast of indptr would corrupt, are hashed in row slices)
//   minhash_cuda_fini(gen)                               python.cc:401-410
// plus minhash_cuda_calc_batch(gen, [matrices], kernel=0), which hashes
// several csr_matrix objects in one batched launch, and
// minhash_cuda_save_vars(gen, path, seed_only=False) /
// minhash_cuda_load_vars(gen, path) for parameter files.
//
// Exception mapping (python.cc:119-142): InvalidArguments and NoSuchDevice ->
// ValueError, MemoryAllocationFailure -> MemoryError, MemoryCopyError ->
// RuntimeError, RuntimeError and unknown codes -> AssertionError.  The GIL is
// released around every engine call.
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#define NPY_NO_DEPRECATED_API NPY_1_7_API_VERSION
#include <numpy/arrayobject.h>

#include <algorithm>
#include <cstdlib>
#include <ctime>
#include <vector>

#include "minhashcuda.h"

namespace {

// Owned reference, released on scope exit.
struct Ref {
  PyObject *p = nullptr;
  Ref() = default;
  explicit Ref(PyObject *o) : p(o) {}
  Ref(const Ref &) = delete;
  Ref(Ref &&o) noexcept : p(o.p) { o.p = nullptr; }
  Ref &operator=(const Ref &) = delete;
  Ref &operator=(Ref &&o) noexcept {
    if (this != &o) {
      Py_XDECREF(p);
      p = o.p;
      o.p = nullptr;
    }
    return *this;
  }
  ~Ref() { Py_XDECREF(p); }
  PyObject *get() const { return p; }
  PyArrayObject *arr() const { return reinterpret_cast<PyArrayObject *>(p); }
  PyObject *release() {
    PyObject *o = p;
    p = nullptr;
    return o;
  }
  explicit operator bool() const { return p != nullptr; }
};

// Sets the Python exception for a failed engine call; returns NULL.
PyObject *raise(MHCUDAResult code, const char *fn) {
  switch (code) {
    case mhcudaInvalidArguments:
      PyErr_Format(PyExc_ValueError, "Invalid arguments were passed to %s", fn);
      break;
    case mhcudaNoSuchDevice:
      PyErr_SetString(PyExc_ValueError, "No such GPU device exists");
      break;
    case mhcudaMemoryAllocationFailure:
      PyErr_SetString(PyExc_MemoryError, "Failed to allocate memory on GPU");
      break;
    case mhcudaMemoryCopyError:
      PyErr_SetString(PyExc_RuntimeError, "hipMemcpy failed");
      break;
    case mhcudaRuntimeError:
      PyErr_Format(PyExc_AssertionError, "%s failure (bug?)", fn);
      break;
    default:
      PyErr_Format(PyExc_AssertionError, "Unknown error code returned from %s", fn);
      break;
  }
  return nullptr;
}

MinhashCudaGenerator *gen_from(unsigned long long ptr) {
  MinhashCudaGenerator *gen =
      reinterpret_cast<MinhashCudaGenerator *>(static_cast<uintptr_t>(ptr));
  if (gen == nullptr) PyErr_SetString(PyExc_ValueError, "MinHashCuda Generator pointer is null.");
  return gen;
}

// A CSR matrix converted to the engine's dtypes (python.cc:295-333).
struct Csr {
  Ref weights, cols, rows;
  uint32_t length = 0;
};

// `obj` as a C-contiguous array of uint32 bit patterns.  int32 and uint32
// arrays are used in place (the force-cast of python.cc:317,323 would produce
// the same bits with a copy); other integer types are force-cast.
Ref as_u32(PyObject *obj) {
  Ref arr(PyArray_FROM_OF(obj, NPY_ARRAY_IN_ARRAY));
  if (arr) {
    const int t = PyArray_TYPE(arr.arr());
    if ((t == NPY_INT32 || t == NPY_UINT32) && PyArray_ITEMSIZE(arr.arr()) == 4) return arr;
  } else {
    PyErr_Clear();
  }
  return Ref(PyArray_FROM_OTF(obj, NPY_UINT32, NPY_ARRAY_FORCECAST | NPY_ARRAY_IN_ARRAY));
}

// python.cc:281-294: only scipy.sparse.csr_matrix instances are accepted.
bool check_csr_type(PyObject *matrix) {
  Ref sparse(PyImport_ImportModule("scipy.sparse"));
  if (!sparse) {
    PyErr_SetString(PyExc_ImportError, "Failed to import scipy.sparse.csr_matrix");
    return false;
  }
  Ref type(PyObject_GetAttrString(sparse.get(), "csr_matrix"));
  if (!type) {
    PyErr_SetString(PyExc_ImportError, "Failed to import scipy.sparse.csr_matrix");
    return false;
  }
  if (!PyObject_TypeCheck(matrix, reinterpret_cast<PyTypeObject *>(type.get()))) {
    PyErr_SetString(PyExc_TypeError, "The second argument must be of type scipy.sparse.csr_matrix");
    return false;
  }
  return true;
}

// data (float32, no FORCECAST: float64 is refused, as in python.cc:311) and
// indices as uint32 bit patterns.
bool load_values(PyObject *matrix, Csr *out) {
  Ref data(PyObject_GetAttrString(matrix, "data"));
  if (!data) return false;
  out->weights = Ref(PyArray_FROM_OTF(data.get(), NPY_FLOAT32, NPY_ARRAY_IN_ARRAY));
  if (!out->weights) {
    PyErr_SetString(PyExc_ValueError, "Failed to convert csr_matrix.data to numpy.array");
    return false;
  }
  Ref indices(PyObject_GetAttrString(matrix, "indices"));
  if (!indices) return false;
  out->cols = as_u32(indices.get());
  if (!out->cols) {
    PyErr_SetString(PyExc_ValueError, "Failed to convert csr_matrix.indices to numpy.array");
    return false;
  }
  return true;
}

// The last indptr entry (the matrix's nonzero count when indptr[0] == 0).
bool last_position(PyObject *matrix, long long *last) {
  Ref indptr(PyObject_GetAttrString(matrix, "indptr"));
  if (!indptr) return false;
  Ref last_obj(PySequence_GetItem(indptr.get(), -1));
  if (!last_obj) return false;
  *last = PyLong_AsLongLong(last_obj.get());
  return !(*last == -1 && PyErr_Occurred());
}

bool load_csr(PyObject *matrix, Csr *out) {
  if (!check_csr_type(matrix) || !load_values(matrix, out)) return false;
  Ref indptr(PyObject_GetAttrString(matrix, "indptr"));
  if (!indptr) return false;
  // the engine's indptr is uint32 (minhashcuda.h:117); minhash_cuda_calc
  // cuts larger matrices into row slices itself, a batch must be given them
  long long last = 0;
  if (!last_position(matrix, &last)) return false;
  if (last > 0xffffffffLL) {
    PyErr_SetString(PyExc_ValueError,
                    "csr_matrix has 2^32 or more nonzeros: hash it with minhash_cuda_calc");
    return false;
  }
  // the positions must be non-decreasing and inside data and indices, checked
  // on their int64 values (a negative position could otherwise wrap into range)
  {
    Ref ip64(PyArray_FROM_OTF(indptr.get(), NPY_INT64, NPY_ARRAY_IN_ARRAY | NPY_ARRAY_FORCECAST));
    if (!ip64) {
      PyErr_SetString(PyExc_ValueError, "Failed to convert csr_matrix.indptr to numpy.array");
      return false;
    }
    const npy_intp n = PyArray_SIZE(ip64.arr());
    const int64_t *ip = static_cast<const int64_t *>(PyArray_DATA(ip64.arr()));
    const int64_t avail =
        std::min<int64_t>(PyArray_SIZE(out->weights.arr()), PyArray_SIZE(out->cols.arr()));
    bool ok = n >= 1 && ip[0] >= 0 && ip[n - 1] <= avail;
    for (npy_intp r = 0; ok && r + 1 < n; r++) ok = ip[r + 1] >= ip[r];
    if (n >= 1 && !ok) {
      PyErr_SetString(PyExc_ValueError,
                      "csr_matrix.indptr must be non-decreasing and within data and indices");
      return false;
    }
  }
  out->rows = as_u32(indptr.get());
  if (!out->rows) {
    PyErr_SetString(PyExc_ValueError, "Failed to convert csr_matrix.indptr to numpy.array");
    return false;
  }
  const npy_intp n = PyArray_SIZE(out->rows.arr());
  if (n < 1) {
    PyErr_SetString(PyExc_ValueError, "csr_matrix.indptr is empty");
    return false;
  }
  if (n - 1 > 0xffffffffLL) {
    PyErr_SetString(PyExc_ValueError, "csr_matrix has 2^32 or more rows");
    return false;
  }
  out->length = static_cast<uint32_t>(n - 1);
  return true;
}

PyObject *new_hashes(uint32_t length, uint16_t samples) {
  npy_intp dims[] = {static_cast<npy_intp>(length), samples, 2};
  return PyArray_EMPTY(3, dims, NPY_UINT32, 0);
}

PyObject *py_init(PyObject *, PyObject *args, PyObject *kwargs) {
  unsigned int dim = 0, seed = static_cast<unsigned int>(time(nullptr)), devices = 0;
  unsigned short samples = 0;
  int deferred = 0, verbosity = 0;
  static const char *kwlist[] = {"dim", "samples", "seed", "deferred", "devices", "verbosity", nullptr};
  if (!PyArg_ParseTupleAndKeywords(args, kwargs, "IH|IpIi", const_cast<char **>(kwlist), &dim,
                                   &samples, &seed, &deferred, &devices, &verbosity))
    return nullptr;
  MHCUDAResult result = mhcudaSuccess;
  MinhashCudaGenerator *gen;
  Py_BEGIN_ALLOW_THREADS
  gen = mhcuda_init(dim, samples, seed, deferred, devices, verbosity, &result);
  Py_END_ALLOW_THREADS
  if (result != mhcudaSuccess || gen == nullptr)
    return raise(result == mhcudaSuccess ? mhcudaRuntimeError : result, "minhash_cuda_init");
  return PyLong_FromUnsignedLongLong(reinterpret_cast<uintptr_t>(gen));
}

PyObject *py_retrieve_vars(PyObject *, PyObject *args) {
  unsigned long long ptr = 0;
  if (!PyArg_ParseTuple(args, "K", &ptr)) return nullptr;
  MinhashCudaGenerator *gen = gen_from(ptr);
  if (!gen) return nullptr;
  const MinhashCudaGeneratorParameters p = mhcuda_get_parameters(gen);
  npy_intp dims[] = {p.samples, static_cast<npy_intp>(p.dim)};
  Ref rs(PyArray_EMPTY(2, dims, NPY_FLOAT32, 0));
  Ref ln_cs(PyArray_EMPTY(2, dims, NPY_FLOAT32, 0));
  Ref betas(PyArray_EMPTY(2, dims, NPY_FLOAT32, 0));
  if (!rs || !ln_cs || !betas) return nullptr;
  float *prs = static_cast<float *>(PyArray_DATA(rs.arr()));
  float *plc = static_cast<float *>(PyArray_DATA(ln_cs.arr()));
  float *pb = static_cast<float *>(PyArray_DATA(betas.arr()));
  MHCUDAResult result;
  Py_BEGIN_ALLOW_THREADS
  result = mhcuda_retrieve_random_vars(gen, prs, plc, pb);
  Py_END_ALLOW_THREADS
  if (result != mhcudaSuccess) return raise(result, "minhash_cuda_retrieve_vars");
  return Py_BuildValue("NNN", rs.release(), ln_cs.release(), betas.release());
}

PyObject *py_assign_vars(PyObject *, PyObject *args) {
  unsigned long long ptr = 0;
  PyObject *objs[3];
  if (!PyArg_ParseTuple(args, "KOOO", &ptr, &objs[0], &objs[1], &objs[2])) return nullptr;
  MinhashCudaGenerator *gen = gen_from(ptr);
  if (!gen) return nullptr;
  const MinhashCudaGeneratorParameters p = mhcuda_get_parameters(gen);
  const int64_t size = (int64_t)p.dim * p.samples;
  static const char *names[3] = {"rs", "ln_cs", "betas"};
  Ref arrs[3];
  const float *data[3];
  for (int i = 0; i < 3; i++) {
    arrs[i] = Ref(PyArray_FROM_OTF(objs[i], NPY_FLOAT32, NPY_ARRAY_IN_ARRAY));
    if (!arrs[i]) {
      PyErr_Format(PyExc_ValueError, "Failed to convert %s to numpy array", names[i]);
      return nullptr;
    }
    if (PyArray_SIZE(arrs[i].arr()) != size) {
      PyErr_Format(PyExc_ValueError, "%s.size must be equal to dim * samples", names[i]);
      return nullptr;
    }
    data[i] = static_cast<const float *>(PyArray_DATA(arrs[i].arr()));
  }
  MHCUDAResult result;
  Py_BEGIN_ALLOW_THREADS
  result = mhcuda_assign_random_vars(gen, data[0], data[1], data[2]);
  Py_END_ALLOW_THREADS
  if (result != mhcudaSuccess) return raise(result, "minhash_cuda_assign_vars");
  Py_RETURN_NONE;
}

// Nonzeros per row slice of a matrix beyond the engine's uint32 indptr
// (MHCUDA_PY_SLICE_NNZ overrides it).
uint64_t slice_nnz() {
  if (const char *env = std::getenv("MHCUDA_PY_SLICE_NNZ")) {
    const unsigned long long v = std::strtoull(env, nullptr, 10);
    if (v > 0 && v <= 0xffffffffull) return v;
  }
  return 1ull << 31;
}

// Rows [row_start, row_finish) of a matrix whose selected nonzeros exceed
// slice_nnz() (in particular 2^32 or more, which the engine's uint32 indptr
// cannot address): consecutive row slices of at most slice_nnz() nonzeros,
// each with its own rebased uint32 indptr, with weights and indices passed in
// place at the slice's offset, all hashed in one pipelined mhcuda_calc_batch
// call.  The hashes are those of a single call over the same rows.
PyObject *calc_sliced(MinhashCudaGenerator *gen, uint32_t row_start, uint32_t row_finish,
                      const Csr &csr, PyArrayObject *indptr) {
  const int64_t *ip = static_cast<const int64_t *>(PyArray_DATA(indptr));
  const uint64_t limit = slice_nnz();
  std::vector<std::vector<uint32_t>> rebased;
  std::vector<MHCUDAMatrix> mats;
  const MinhashCudaGeneratorParameters p = mhcuda_get_parameters(gen);
  Ref out(new_hashes(row_finish - row_start, p.samples));
  if (!out) return nullptr;
  uint32_t *output = static_cast<uint32_t *>(PyArray_DATA(out.arr()));
  const float *weights = static_cast<const float *>(PyArray_DATA(csr.weights.arr()));
  const uint32_t *cols = static_cast<const uint32_t *>(PyArray_DATA(csr.cols.arr()));
  for (uint32_t r = row_start; r < row_finish;) {
    const int64_t base = ip[r];
    uint32_t e = r + 1;
    while (e < row_finish && (uint64_t)(ip[e + 1] - base) <= limit) e++;
    if ((uint64_t)(ip[e] - base) > 0xffffffffull) {
      PyErr_Format(PyExc_ValueError, "row %u has 2^32 or more nonzeros", r);
      return nullptr;
    }
    std::vector<uint32_t> rows(e - r + 1);
    for (uint32_t i = r; i <= e; i++) rows[i - r] = (uint32_t)(ip[i] - base);
    rebased.push_back(std::move(rows));
    mats.push_back(MHCUDAMatrix{weights + base, cols + base, rebased.back().data(), e - r,
                                output + (uint64_t)(r - row_start) * p.samples * 2u});
    r = e;
  }
  // the vectors moved into `rebased` keep their buffers: fix the pointers up
  for (size_t i = 0; i < mats.size(); i++) mats[i].rows = rebased[i].data();
  MHCUDAResult result;
  Py_BEGIN_ALLOW_THREADS
  result = mhcuda_calc_batch(gen, static_cast<uint32_t>(mats.size()), mats.data(),
                             mhcudaKernelAuto);
  Py_END_ALLOW_THREADS
  if (result != mhcudaSuccess) return raise(result, "minhash_cuda_calc");
  return out.release();
}

PyObject *py_calc(PyObject *, PyObject *args, PyObject *kwargs) {
  unsigned long long ptr = 0;
  PyObject *matrix = nullptr;
  unsigned int row_start = 0, row_finish = 0xffffffffu;
  static const char *kwlist[] = {"gen", "matrix", "row_start", "row_finish", nullptr};
  if (!PyArg_ParseTupleAndKeywords(args, kwargs, "KO|II", const_cast<char **>(kwlist), &ptr,
                                   &matrix, &row_start, &row_finish))
    return nullptr;
  MinhashCudaGenerator *gen = gen_from(ptr);
  if (!gen) return nullptr;
  if (!check_csr_type(matrix)) return nullptr;
  Csr csr;
  if (!load_values(matrix, &csr)) return nullptr;
  Ref indptr_obj(PyObject_GetAttrString(matrix, "indptr"));
  if (!indptr_obj) return nullptr;
  // positions as int64 (a copy of rows + 1 values unless already int64)
  Ref indptr(PyArray_FROM_OTF(indptr_obj.get(), NPY_INT64, NPY_ARRAY_IN_ARRAY | NPY_ARRAY_FORCECAST));
  if (!indptr) {
    PyErr_SetString(PyExc_ValueError, "Failed to convert csr_matrix.indptr to numpy.array");
    return nullptr;
  }
  const npy_intp n = PyArray_SIZE(indptr.arr());
  if (n < 1) {
    PyErr_SetString(PyExc_ValueError, "csr_matrix.indptr is empty");
    return nullptr;
  }
  if (n - 1 > 0xffffffffLL) {
    PyErr_SetString(PyExc_ValueError, "csr_matrix has 2^32 or more rows");
    return nullptr;
  }
  uint32_t length = static_cast<uint32_t>(n - 1);
  // row slicing, python.cc:335-359
  if (row_start >= length) {
    PyErr_SetString(PyExc_ValueError,
                    "minhash_cuda_calc: row_start must be less than the total number of rows "
                    "in the input matrix.");
    return nullptr;
  }
  row_finish = std::min<uint32_t>(row_finish, length);
  if (row_finish <= row_start) {
    PyErr_SetString(PyExc_ValueError, "minhash_cuda_calc: row_finish must be greater than row_start.");
    return nullptr;
  }
  // the selected rows must address positions inside data and indices
  const int64_t *ip = static_cast<const int64_t *>(PyArray_DATA(indptr.arr()));
  const int64_t avail = std::min<int64_t>(PyArray_SIZE(csr.weights.arr()), PyArray_SIZE(csr.cols.arr()));
  for (uint32_t r = row_start; r < row_finish; r++) {
    if (ip[r] < 0 || ip[r + 1] < ip[r] || ip[r + 1] > avail) {
      PyErr_SetString(PyExc_ValueError,
                      "csr_matrix.indptr must be non-decreasing and within data and indices");
      return nullptr;
    }
  }
  if (ip[length] > 0xffffffffLL || (uint64_t)(ip[row_finish] - ip[row_start]) > slice_nnz())
    return calc_sliced(gen, row_start, row_finish, csr, indptr.arr());
  // the whole matrix fits the uint32 indptr: it stays absolute, so the row
  // slice is zero-copy (no rebased indptr is built)
  csr.rows = as_u32(indptr_obj.get());
  if (!csr.rows) {
    PyErr_SetString(PyExc_ValueError, "Failed to convert csr_matrix.indptr to numpy.array");
    return nullptr;
  }
  length = row_finish - row_start;
  const float *weights = static_cast<const float *>(PyArray_DATA(csr.weights.arr()));
  const uint32_t *cols = static_cast<const uint32_t *>(PyArray_DATA(csr.cols.arr()));
  const uint32_t *rows = static_cast<const uint32_t *>(PyArray_DATA(csr.rows.arr())) + row_start;
  const MinhashCudaGeneratorParameters p = mhcuda_get_parameters(gen);
  Ref out(new_hashes(length, p.samples));
  if (!out) return nullptr;
  uint32_t *output = static_cast<uint32_t *>(PyArray_DATA(out.arr()));
  MHCUDAResult result;
  Py_BEGIN_ALLOW_THREADS
  result = mhcuda_calc(gen, weights, cols, rows, length, output);
  Py_END_ALLOW_THREADS
  if (result != mhcudaSuccess) return raise(result, "minhash_cuda_calc");
  return out.release();
}

PyObject *py_calc_batch(PyObject *, PyObject *args, PyObject *kwargs) {
  unsigned long long ptr = 0;
  PyObject *seq = nullptr;
  int kernel = mhcudaKernelAuto;
  static const char *kwlist[] = {"gen", "matrices", "kernel", nullptr};
  if (!PyArg_ParseTupleAndKeywords(args, kwargs, "KO|i", const_cast<char **>(kwlist), &ptr, &seq,
                                   &kernel))
    return nullptr;
  MinhashCudaGenerator *gen = gen_from(ptr);
  if (!gen) return nullptr;
  Ref fast(PySequence_Fast(seq, "matrices must be a sequence of csr_matrix"));
  if (!fast) return nullptr;
  const Py_ssize_t count = PySequence_Fast_GET_SIZE(fast.get());
  if (count == 0) return PyList_New(0);
  // every matrix is checked before the generator is used
  std::vector<Csr> csrs(count);
  for (Py_ssize_t i = 0; i < count; i++)
    if (!load_csr(PySequence_Fast_GET_ITEM(fast.get(), i), &csrs[i])) return nullptr;
  const MinhashCudaGeneratorParameters p = mhcuda_get_parameters(gen);
  std::vector<MHCUDAMatrix> mats(count);
  Ref outs(PyList_New(count));
  if (!outs) return nullptr;
  for (Py_ssize_t i = 0; i < count; i++) {
    PyObject *o = new_hashes(csrs[i].length, p.samples);
    if (!o) return nullptr;
    PyList_SET_ITEM(outs.get(), i, o);
    mats[i].weights = static_cast<const float *>(PyArray_DATA(csrs[i].weights.arr()));
    mats[i].cols = static_cast<const uint32_t *>(PyArray_DATA(csrs[i].cols.arr()));
    mats[i].rows = static_cast<const uint32_t *>(PyArray_DATA(csrs[i].rows.arr()));
    mats[i].length = csrs[i].length;
    mats[i].output = static_cast<uint32_t *>(PyArray_DATA(reinterpret_cast<PyArrayObject *>(o)));
  }
  MHCUDAResult result;
  Py_BEGIN_ALLOW_THREADS
  result = mhcuda_calc_batch(gen, static_cast<uint32_t>(count), mats.data(), kernel);
  Py_END_ALLOW_THREADS
  if (result != mhcudaSuccess) return raise(result, "minhash_cuda_calc_batch");
  return outs.release();
}

// Parameter files (see mhcuda_save_random_vars): ValueError for a malformed or
// mismatched file, OSError when the file cannot be read or written.
PyObject *raise_file(MHCUDAResult code, const char *fn, const char *path) {
  if (code == mhcudaInvalidArguments) {
    PyErr_Format(PyExc_ValueError,
                 "%s: %s is not a valid parameter file for this generator "
                 "(or the generator holds no tables)", fn, path);
    return nullptr;
  }
  if (code == mhcudaMemoryCopyError) {
    PyErr_Format(PyExc_OSError, "%s: failed to read or write %s", fn, path);
    return nullptr;
  }
  return raise(code, fn);
}

PyObject *py_save_vars(PyObject *, PyObject *args, PyObject *kwargs) {
  unsigned long long ptr = 0;
  PyObject *path_obj = nullptr;
  int seed_only = 0;
  static const char *kwlist[] = {"gen", "path", "seed_only", nullptr};
  if (!PyArg_ParseTupleAndKeywords(args, kwargs, "KO&|p", const_cast<char **>(kwlist), &ptr,
                                   PyUnicode_FSConverter, &path_obj, &seed_only))
    return nullptr;
  Ref path(path_obj);
  MinhashCudaGenerator *gen = gen_from(ptr);
  if (!gen) return nullptr;
  const char *p = PyBytes_AS_STRING(path.get());
  MHCUDAResult result;
  Py_BEGIN_ALLOW_THREADS
  result = mhcuda_save_random_vars(gen, p, seed_only);
  Py_END_ALLOW_THREADS
  if (result != mhcudaSuccess) return raise_file(result, "minhash_cuda_save_vars", p);
  Py_RETURN_NONE;
}

PyObject *py_load_vars(PyObject *, PyObject *args, PyObject *kwargs) {
  unsigned long long ptr = 0;
  PyObject *path_obj = nullptr;
  static const char *kwlist[] = {"gen", "path", nullptr};
  if (!PyArg_ParseTupleAndKeywords(args, kwargs, "KO&", const_cast<char **>(kwlist), &ptr,
                                   PyUnicode_FSConverter, &path_obj))
    return nullptr;
  Ref path(path_obj);
  MinhashCudaGenerator *gen = gen_from(ptr);
  if (!gen) return nullptr;
  const char *p = PyBytes_AS_STRING(path.get());
  MHCUDAResult result;
  Py_BEGIN_ALLOW_THREADS
  result = mhcuda_load_random_vars(gen, p);
  Py_END_ALLOW_THREADS
  if (result != mhcudaSuccess) return raise_file(result, "minhash_cuda_load_vars", p);
  Py_RETURN_NONE;
}

PyObject *py_fini(PyObject *, PyObject *args) {
  unsigned long long ptr = 0;
  if (!PyArg_ParseTuple(args, "K", &ptr)) return nullptr;
  MinhashCudaGenerator *gen = reinterpret_cast<MinhashCudaGenerator *>(static_cast<uintptr_t>(ptr));
  Py_BEGIN_ALLOW_THREADS
  mhcuda_fini(gen);
  Py_END_ALLOW_THREADS
  Py_RETURN_NONE;
}

PyMethodDef methods[] = {
    {"minhash_cuda_init", reinterpret_cast<PyCFunction>(reinterpret_cast<void (*)(void)>(py_init)),
     METH_VARARGS | METH_KEYWORDS, "Prepares Weighted MinHash internal state on the GPU."},
    {"minhash_cuda_retrieve_vars", py_retrieve_vars, METH_VARARGS,
     "Copies the random variables rs, ln_cs and betas from the generator to the host."},
    {"minhash_cuda_assign_vars", py_assign_vars, METH_VARARGS,
     "Assigns the random variables rs, ln_cs and betas to the generator."},
    {"minhash_cuda_calc", reinterpret_cast<PyCFunction>(reinterpret_cast<void (*)(void)>(py_calc)),
     METH_VARARGS | METH_KEYWORDS, "Calculates Weighted MinHashes."},
    {"minhash_cuda_calc_batch",
     reinterpret_cast<PyCFunction>(reinterpret_cast<void (*)(void)>(py_calc_batch)),
     METH_VARARGS | METH_KEYWORDS,
     "Calculates Weighted MinHashes of several csr_matrix objects in one batch."},
    {"minhash_cuda_save_vars",
     reinterpret_cast<PyCFunction>(reinterpret_cast<void (*)(void)>(py_save_vars)),
     METH_VARARGS | METH_KEYWORDS,
     "Writes the generator's random variables to a parameter file (seed_only: just the "
     "seed when the tables were generated on the device)."},
    {"minhash_cuda_load_vars",
     reinterpret_cast<PyCFunction>(reinterpret_cast<void (*)(void)>(py_load_vars)),
     METH_VARARGS | METH_KEYWORDS,
     "Loads a parameter file into a generator of the same dim and samples."},
    {"minhash_cuda_fini", py_fini, METH_VARARGS,
     "Disposes Weighted MinHash internal state on the GPU."},
    {nullptr, nullptr, 0, nullptr}};

PyModuleDef moduledef = {PyModuleDef_HEAD_INIT,
                         "libMHCUDA",
                         "Weighted MinHash on AMD MI355X (gfx950) GPUs.",
                         -1,
                         methods,
                         nullptr,
                         nullptr,
                         nullptr,
                         nullptr};

}  // namespace

PyMODINIT_FUNC PyInit_libMHCUDA(void) {
  PyObject *m = PyModule_Create(&moduledef);
  if (m == nullptr) return nullptr;
  import_array();
  return m;
}
