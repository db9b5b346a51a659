// Python module cholespy_amd._cholespy_core: MatrixType, CholeskySolverF,
// CholeskySolverD and __version__, with the reference's argument names,
// validation order, error types and messages (src/main.cpp:16-136,
// src/docstr.h:1-48).
//
// nanobind is not available, so arrays are taken through the DLPack protocol
// (__dlpack__) directly: NumPy, PyTorch CPU and PyTorch ROCm tensors work
// unchanged. Constructor arrays may be any integer / float dtype and any
// stride (the reference's implicit conversion of `ii`, `jj`, `x`, e.g. int64
// index arrays); solve() takes b and x without conversion (src/main.cpp:108-109).
// A ROCm tensor reports DLPack device type 10 (kDLROCM); it is accepted
// where the reference accepts CUDA (2).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "cholesky_solver.hpp"
#include "dlpack_abi.hpp"
#include "docstrings.hpp"

namespace py = pybind11;
using namespace cholespy_amd;

namespace {

enum class Dev { CPU, GPU, Other };

// A borrowed view of an array exported through DLPack. Calls the producer's
// deleter when it goes out of scope.
class DLView {
public:
    explicit DLView(py::handle obj) {
        if (!py::hasattr(obj, "__dlpack__"))
            throw py::type_error("Expected an array supporting the DLPack protocol (__dlpack__).");
        py::object cap;
        bool versioned = false;
        try {
            cap = obj.attr("__dlpack__")();
        } catch (py::error_already_set &e) {
            // NumPy refuses legacy export of read-only arrays; DLPack 1.0 has a
            // read-only flag and is allowed.
            if (!e.matches(PyExc_BufferError)) throw;
            cap = obj.attr("__dlpack__")(py::arg("max_version") = py::make_tuple(1, 0));
            versioned = true;
        }
        PyObject *c = cap.ptr();
        if (PyCapsule_IsValid(c, "dltensor_versioned")) versioned = true;
        if (versioned) {
            vmt_ = static_cast<dl::ManagedTensorVersioned *>(PyCapsule_GetPointer(c, "dltensor_versioned"));
            if (!vmt_) throw py::error_already_set();
            PyCapsule_SetName(c, "used_dltensor_versioned");
            t_ = &vmt_->dl_tensor;
            readonly_ = (vmt_->flags & 1u) != 0;
        } else {
            mt_ = static_cast<dl::ManagedTensor *>(PyCapsule_GetPointer(c, "dltensor"));
            if (!mt_) throw py::error_already_set();
            PyCapsule_SetName(c, "used_dltensor");
            t_ = &mt_->dl_tensor;
        }
    }
    ~DLView() {
        if (mt_ && mt_->deleter) mt_->deleter(mt_);
        if (vmt_ && vmt_->deleter) vmt_->deleter(vmt_);
    }
    DLView(const DLView &) = delete;
    DLView &operator=(const DLView &) = delete;

    Dev device() const {
        const int dt = t_->device.device_type;
        if (dt == dl::kCPU) return Dev::CPU;
        if (dt == dl::kCUDA || dt == dl::kROCM) return Dev::GPU;
        return Dev::Other;
    }
    int device_type() const { return t_->device.device_type; }
    int ndim() const { return t_->ndim; }
    int64_t shape(int i) const { return t_->shape[i]; }
    int64_t stride(int i) const { return t_->strides ? t_->strides[i] : contiguous_stride(i); }
    void *data() const { return static_cast<char *>(t_->data) + t_->byte_offset; }
    const dl::DataType &dtype() const { return t_->dtype; }
    bool readonly() const { return readonly_; }
    bool c_contig() const {
        if (!t_->strides) return true;
        int64_t expect = 1;
        for (int i = t_->ndim - 1; i >= 0; --i) {
            if (t_->shape[i] != 1 && t_->strides[i] != expect) return false;
            expect *= t_->shape[i];
        }
        return true;
    }
    int64_t size() const {
        int64_t s = 1;
        for (int i = 0; i < t_->ndim; ++i) s *= t_->shape[i];
        return s;
    }

private:
    int64_t contiguous_stride(int i) const {
        int64_t s = 1;
        for (int j = t_->ndim - 1; j > i; --j) s *= t_->shape[j];
        return s;
    }
    dl::ManagedTensor *mt_ = nullptr;
    dl::ManagedTensorVersioned *vmt_ = nullptr;
    dl::Tensor *t_ = nullptr;
    bool readonly_ = false;
};

// Ensures obj exposes __dlpack__; plain sequences go through numpy.asarray.
py::object as_exportable(py::handle obj) {
    if (py::hasattr(obj, "__dlpack__")) return py::reinterpret_borrow<py::object>(obj);
    return py::module_::import("numpy").attr("asarray")(obj);
}

// Plan arrays of a GPU handle (cholespy_solver_plan_arrays) as a dict of numpy arrays.
py::dict plan_arrays(const cholespy_gpu_solver *h) {
    const char *names[] = {"start", "rptr", "level_fwd", "level_bwd", "fwd_task", "bwd_task", "fwd_task_wide",
                           "bwd_task_wide"};
    py::dict d;
    for (int w = 0; w < 8; ++w) {
        int64_t size = 0;
        throw_on_error(cholespy_solver_plan_arrays(h, w, nullptr, 0, &size));
        py::array_t<int32_t> a(size);
        throw_on_error(cholespy_solver_plan_arrays(h, w, a.mutable_data(), size, nullptr));
        if (w >= 4) a.resize(std::vector<py::ssize_t>{size / CHOLESPY_TASK_INTS, CHOLESPY_TASK_INTS});
        d[names[w]] = a;
    }
    return d;
}

template <typename T>
T read_elem(const char *p, const dl::DataType &dt) {
    switch (dt.code) {
    case dl::kInt:
        switch (dt.bits) {
        case 8: { int8_t v; std::memcpy(&v, p, 1); return (T)v; }
        case 16: { int16_t v; std::memcpy(&v, p, 2); return (T)v; }
        case 32: { int32_t v; std::memcpy(&v, p, 4); return (T)v; }
        case 64: { int64_t v; std::memcpy(&v, p, 8); return (T)v; }
        }
        break;
    case dl::kUInt:
        switch (dt.bits) {
        case 8: { uint8_t v; std::memcpy(&v, p, 1); return (T)v; }
        case 16: { uint16_t v; std::memcpy(&v, p, 2); return (T)v; }
        case 32: { uint32_t v; std::memcpy(&v, p, 4); return (T)v; }
        case 64: { uint64_t v; std::memcpy(&v, p, 8); return (T)v; }
        }
        break;
    case dl::kFloat:
        if (std::is_integral<T>::value) break;   // no implicit float -> index conversion
        switch (dt.bits) {
        case 32: { float v; std::memcpy(&v, p, 4); return (T)v; }
        case 64: { double v; std::memcpy(&v, p, 8); return (T)v; }
        }
        break;
    }
    throw py::type_error("Unsupported array dtype for this argument.");
}

// Copies a 1-D array (CPU or GPU, any stride) into a host vector of T,
// converting the element type: the constructor's implicit conversion.
template <typename T>
std::vector<T> to_host(const DLView &v) {
    const int64_t n = v.shape(0);
    const int64_t st = v.stride(0);
    const int eb = v.dtype().bits / 8;
    if (v.dtype().lanes != 1 || eb == 0) throw py::type_error("Unsupported array dtype.");
    std::vector<T> out((size_t)n);
    if (n == 0) return out;
    const char *base = static_cast<const char *>(v.data());
    std::vector<char> staged;
    if (v.device() == Dev::GPU) {
        const int64_t lo = st >= 0 ? 0 : (n - 1) * st;
        const int64_t span = ((n - 1) * (st >= 0 ? st : -st) + 1) * eb;
        staged.resize((size_t)span);
        throw_on_error(cholespy_copy_to_host(staged.data(), base + lo * eb, (size_t)span));
        base = staged.data() - lo * eb;
    }
    for (int64_t i = 0; i < n; ++i) out[(size_t)i] = read_elem<T>(base + i * st * eb, v.dtype());
    return out;
}

// Whether the constructor's GPU arrays can go to cholespy_normalize_input
// as they are: one-dimensional and contiguous, indices int32 or int64 (the
// same for both), values float32 or float64.
bool device_normalizable(const DLView &ii, const DLView &jj, const DLView &x) {
    auto contiguous = [](const DLView &v) { return v.ndim() == 1 && (v.shape(0) <= 1 || v.stride(0) == 1); };
    const dl::DataType &ti = ii.dtype(), &tj = jj.dtype(), &tx = x.dtype();
    const bool idx_ok = ti.code == dl::kInt && ti.lanes == 1 && (ti.bits == 32 || ti.bits == 64) &&
                        tj.code == ti.code && tj.bits == ti.bits && tj.lanes == 1;
    const bool val_ok = tx.code == dl::kFloat && tx.lanes == 1 && (tx.bits == 32 || tx.bits == 64);
    return idx_ok && val_ok && contiguous(ii) && contiguous(jj) && contiguous(x);
}

template <typename Float>
void declare_solver(py::module_ &m, const char *name, const char *doc) {
    using Class = CholeskySolver<Float>;
    py::class_<Class>(m, name, doc)
        .def(py::init([](uint32_t n_rows, py::object ii_o, py::object jj_o, py::object x_o,
                         MatrixType type) {
                 ii_o = as_exportable(ii_o);
                 jj_o = as_exportable(jj_o);
                 x_o = as_exportable(x_o);
                 DLView ii(ii_o), jj(jj_o), x(x_o);
                 if (ii.ndim() != 1 || jj.ndim() != 1 || x.ndim() != 1)
                     throw py::type_error("ii, jj and x must be one-dimensional arrays.");
                 // src/main.cpp:28-43
                 if (type == MatrixType::COO) {
                     if (ii.shape(0) != jj.shape(0))
                         throw std::invalid_argument("Sparse COO matrix: the two index arrays should have the same size.");
                     if (ii.shape(0) != x.shape(0))
                         throw std::invalid_argument("Sparse COO matrix: the index and data arrays should have the same size.");
                 } else if (type == MatrixType::CSR) {
                     if (jj.shape(0) != x.shape(0))
                         throw std::invalid_argument("Sparse CSR matrix: the column index and data arrays should have the same size.");
                     if (ii.shape(0) != (int64_t)n_rows + 1)
                         throw std::invalid_argument("Sparse CSR matrix: Invalid size for row pointer array.");
                 } else {
                     if (jj.shape(0) != x.shape(0))
                         throw std::invalid_argument("Sparse CSC matrix: the row index and data arrays should have the same size.");
                     if (ii.shape(0) != (int64_t)n_rows + 1)
                         throw std::invalid_argument("Sparse CSC matrix: Invalid size for column pointer array.");
                 }
                 // src/main.cpp:44-45
                 if (ii.device_type() != jj.device_type() || ii.device_type() != x.device_type())
                     throw std::invalid_argument("All input tensors should be on the same device!");
                 if (n_rows > 0x7fffffffu) throw std::invalid_argument("n_rows does not fit in 32 bits.");
                 const Dev dev = ii.device();
                 if (dev != Dev::CPU && dev != Dev::GPU)
                     throw std::invalid_argument("Unsupported input device! Only CPU and CUDA arrays are supported.");
                 if (dev == Dev::GPU) throw_on_error(cholespy_init());   // src/main.cpp:51
                 // GPU arrays that are contiguous, with 32- or 64-bit integer
                 // indices and float values, are normalised on the device
                 // (only the lower CSC comes to the host); others take the
                 // reference's route of copying the raw arrays first.
                 if (dev == Dev::GPU && device_normalizable(ii, jj, x)) {
                     const int64_t nnz = x.shape(0);
                     LowerCSC A;
                     A.n = (int)n_rows;
                     A.colptr.resize((size_t)n_rows + 1);
                     A.rowidx.resize((size_t)nnz);
                     A.val.resize((size_t)nnz);
                     std::vector<int64_t> edst((size_t)nnz);
                     int64_t m = 0;
                     py::gil_scoped_release release;
                     throw_on_error(cholespy_normalize_input((int)n_rows, nnz, ii.data(), jj.data(), ii.dtype().bits / 8,
                                                             x.data(), x.dtype().bits / 8, (int)type, A.colptr.data(),
                                                             A.rowidx.data(), A.val.data(), edst.data(), &m));
                     A.rowidx.resize((size_t)m);
                     A.val.resize((size_t)m);
                     return std::make_unique<Class>(std::move(A), std::move(edst));
                 }
                 std::vector<int> iv = to_host<int>(ii), jv = to_host<int>(jj);
                 std::vector<double> xv = to_host<double>(x);
                 py::gil_scoped_release release;
                 return std::make_unique<Class>((int)n_rows, (int64_t)xv.size(), iv.data(), jv.data(),
                                                xv.data(), type, dev == Dev::CPU);
             }),
             py::arg("n_rows"), py::arg("ii"), py::arg("jj"), py::arg("x"), py::arg("type"),
             doc_constructor)
        .def(
            "solve",
            [](Class &self, py::object b_o, py::object x_o, uintptr_t stream) {
                DLView b(b_o), x(x_o);
                const bool is_f32 = std::is_same<Float, float>::value;
                const auto good = [&](const DLView &v) {
                    return v.dtype().code == dl::kFloat && v.dtype().bits == (is_f32 ? 32 : 64) &&
                           v.dtype().lanes == 1 && v.c_contig();
                };
                // noconvert (src/main.cpp:108-109): wrong dtype or layout is a type error.
                if (!good(b) || !good(x))
                    throw py::type_error(std::string("solve(): b and x must be C-contiguous ") +
                                         (is_f32 ? "float32" : "float64") +
                                         " arrays (no implicit conversion, x is written in place).");
                if (x.readonly()) throw py::type_error("solve(): x must be writable.");
                // src/main.cpp:83-88
                if (b.ndim() != 1 && b.ndim() != 2)
                    throw std::invalid_argument("Expected 1D or 2D tensors as input.");
                if (b.ndim() != x.ndim() || b.shape(0) != x.shape(0) ||
                    (b.ndim() == 2 && b.shape(1) != x.shape(1)))
                    throw std::invalid_argument("x and b should have the same dimensions.");
                if (b.device() != x.device())
                    throw std::invalid_argument("x and b should be on the same device.");
                const int n_rhs = b.ndim() == 2 ? (int)b.shape(1) : 1;
                // src/main.cpp:91-106
                if (b.device() == Dev::CPU) {
                    if (!self.is_cpu())
                        throw std::invalid_argument("Input device is CPU but the solver was initialized for CUDA.");
                } else if (b.device() == Dev::GPU) {
                    if (self.is_cpu())
                        throw std::invalid_argument("Input device is CUDA but the solver was initialized for CPU.");
                } else {
                    throw std::invalid_argument("Unsupported input device! Only CPU and CUDA arrays are supported.");
                }
                // Not checked by the reference, whose kernels would then read
                // and write out of bounds.
                if (b.shape(0) != self.n())
                    throw std::invalid_argument("b and x should have n_rows rows.");
                if (b.device() == Dev::CPU) {
                    const Float *bp = static_cast<const Float *>(b.data());
                    Float *xp = static_cast<Float *>(x.data());
                    py::gil_scoped_release release;
                    self.solve_cpu(n_rhs, bp, xp);
                } else {
                    self.solve_cuda(n_rhs, b.data(), x.data(), reinterpret_cast<void *>(stream));
                }
            },
            py::arg("b"), py::arg("x"), py::arg("stream") = 0, doc_solve)
        .def(
            "refactorize",
            [](Class &self, py::object x_o) {
                x_o = as_exportable(x_o);
                DLView x(x_o);
                if (x.ndim() != 1) throw py::type_error("x must be a one-dimensional array.");
                if (x.device() != Dev::CPU && x.device() != Dev::GPU)
                    throw std::invalid_argument("Unsupported input device! Only CPU and CUDA arrays are supported.");
                if (x.shape(0) != self.input_nnz())
                    throw std::invalid_argument("refactorize(): x must have as many entries as at construction.");
                std::vector<double> xv = to_host<double>(x);
                py::gil_scoped_release release;
                self.refactorize((int64_t)xv.size(), xv.data());
            },
            py::arg("x"), doc_refactorize)
        // Inspection hooks (not part of the reference API): the host factor,
        // the GPU level analysis and the device watchdog status.
        .def("_factor",
             [](const Class &self) {
                 const Factor &f = self.factor();
                 py::dict d;
                 d["perm"] = py::array_t<int>(f.perm.size(), f.perm.data());
                 d["colptr"] = py::array_t<int>(f.colptr.size(), f.colptr.data());
                 d["rowidx"] = py::array_t<int>(f.rowidx.size(), f.rowidx.data());
                 d["val"] = py::array_t<double>(f.val.size(), f.val.data());
                 d["parent"] = py::array_t<int>(f.parent.size(), f.parent.data());
                 return d;
             })
        .def("_matrix",
             [](const Class &self) {
                 const LowerCSC &a = self.matrix();
                 py::dict d;
                 d["colptr"] = py::array_t<int>(a.colptr.size(), a.colptr.data());
                 d["rowidx"] = py::array_t<int>(a.rowidx.size(), a.rowidx.data());
                 d["val"] = py::array_t<double>(a.val.size(), a.val.data());
                 return d;
             })
        .def("_levels",
             [](const Class &self, bool lower) {
                 std::vector<int> order, row_level;
                 self.levels(lower, order, row_level);
                 return py::make_tuple(py::array_t<int>(order.size(), order.data()),
                                       py::array_t<int>(row_level.size(), row_level.data()));
             },
             py::arg("lower"))
        .def("_n_levels", [](const Class &self, bool lower) { return self.n_levels(lower); },
             py::arg("lower"))
        .def("_check", [](const Class &self) { self.check(); })
        .def(
            "autotune",
            [](Class &self, int n_rhs, py::object cache, bool force, bool write) {
                const std::string path = cache.is_none() ? std::string() : py::str(cache).cast<std::string>();
                std::string report;
                {
                    py::gil_scoped_release release;
                    report = self.autotune(n_rhs, path, (force ? 1 : 0) | (write ? 0 : 2));
                }
                // "tuned|cached key=value ..." -> dict
                py::dict d;
                size_t pos = 0;
                bool first = true;
                while (pos < report.size()) {
                    size_t end = report.find(' ', pos);
                    if (end == std::string::npos) end = report.size();
                    const std::string tok = report.substr(pos, end - pos);
                    pos = end + 1;
                    if (first) {
                        d["source"] = tok;
                        first = false;
                        continue;
                    }
                    const size_t eq = tok.find('=');
                    if (eq == std::string::npos) continue;
                    const std::string key = tok.substr(0, eq), val = tok.substr(eq + 1);
                    if (key == "key") d[key.c_str()] = val;
                    else if (key == "ms" || (key.size() > 3 && key.compare(key.size() - 3, 3, "_ms") == 0))
                        d[key.c_str()] = std::stod(val);
                    else d[key.c_str()] = std::stoi(val);
                }
                return d;
            },
            py::arg("n_rhs"), py::arg("cache") = py::none(), py::arg("force") = false, py::arg("write") = true,
            "Times candidate solve configurations for n_rhs right-hand sides on the GPU and keeps the\n"
            "fastest (schedule, tile, block-plan shape, leaf phase, ticket queues); the choice is\n"
            "cached on disk per factor pattern, precision, right-hand-side class and device.\n"
            "Returns the chosen configuration and its time per solve. GPU solvers only.")
        // Watchdog test hooks (cholespy_solver_set_spin_limit / _raise_abort).
        .def("_set_spin_limit", [](Class &self, uint32_t limit) { throw_on_error(cholespy_solver_set_spin_limit(self.gpu(), limit)); },
             py::arg("limit"))
        .def("_raise_abort", [](Class &self) { throw_on_error(cholespy_solver_raise_abort(self.gpu())); })
        .def("_set_trace", [](Class &self, bool on) { throw_on_error(cholespy_solver_set_trace(self.gpu(), on ? 1 : 0)); },
             py::arg("on"))
        .def("_plan_arrays", [](const Class &self) { return plan_arrays(self.gpu()); })
        .def("_trace",
             [](Class &self, bool forward) {
                 const std::vector<int64_t> info = self.plan_info();
                 const int64_t words = 4 * (forward ? std::max(info[4], info[8]) : std::max(info[5], info[9]));
                 py::array_t<uint64_t> out(words);
                 throw_on_error(cholespy_solver_get_trace(self.gpu(), forward ? 1 : 0, out.mutable_data(), words));
                 return out;
             },
             py::arg("forward"))
        .def("_set_path", [](Class &self, int path) { self.set_path(path); }, py::arg("path"))
        .def(
            "_solve_pass",
            [](Class &self, int pass, py::object b_o, py::object x_o) {
                DLView b(b_o), x(x_o);
                if (b.device() != Dev::GPU || x.device() != Dev::GPU || !b.c_contig() || !x.c_contig() ||
                    b.dtype().code != dl::kFloat || b.dtype().bits != 8 * sizeof(Float) ||
                    x.dtype().code != dl::kFloat || x.dtype().bits != 8 * sizeof(Float))
                    throw py::type_error("_solve_pass(): b and x must be C-contiguous GPU arrays of the solver dtype.");
                if (b.ndim() != x.ndim() || b.shape(0) != x.shape(0) || b.shape(0) != self.n() ||
                    (b.ndim() == 2 && b.shape(1) != x.shape(1)))
                    throw std::invalid_argument("x and b should have the same dimensions.");
                self.solve_pass(pass, b.ndim() == 2 ? (int)b.shape(1) : 1, b.data(), x.data());
            },
            py::arg("pass_"), py::arg("b"), py::arg("x"))
        .def(
            "_solve_refined",
            [](Class &self, py::object b_o, py::object x_o, int iterations) {
                DLView b(b_o), x(x_o);
                const auto good = [](const DLView &v) {
                    return v.dtype().code == dl::kFloat && v.dtype().bits == 64 && v.dtype().lanes == 1 &&
                           v.c_contig() && v.device() == Dev::GPU;
                };
                if (!good(b) || !good(x) || x.readonly())
                    throw py::type_error("_solve_refined(): b and x must be C-contiguous float64 GPU arrays.");
                if (b.ndim() != 1 && b.ndim() != 2)
                    throw std::invalid_argument("Expected 1D or 2D tensors as input.");
                if (b.ndim() != x.ndim() || b.shape(0) != x.shape(0) ||
                    (b.ndim() == 2 && b.shape(1) != x.shape(1)))
                    throw std::invalid_argument("x and b should have the same dimensions.");
                if (b.shape(0) != self.n()) throw std::invalid_argument("b and x should have n_rows rows.");
                self.solve_refined(b.ndim() == 2 ? (int)b.shape(1) : 1, static_cast<const double *>(b.data()),
                                   static_cast<double *>(x.data()), iterations);
            },
            py::arg("b"), py::arg("x"), py::arg("iterations") = 2)
        .def(
            "_solve_refined_tol",
            [](Class &self, py::object b_o, py::object x_o, double tol, int max_steps) {
                DLView b(b_o), x(x_o);
                const auto good = [](const DLView &v) {
                    return v.dtype().code == dl::kFloat && v.dtype().bits == 64 && v.dtype().lanes == 1 &&
                           v.c_contig() && v.device() == Dev::GPU;
                };
                if (!good(b) || !good(x) || x.readonly())
                    throw py::type_error("_solve_refined_tol(): b and x must be C-contiguous float64 GPU arrays.");
                if (b.ndim() != 1 && b.ndim() != 2)
                    throw std::invalid_argument("Expected 1D or 2D tensors as input.");
                if (b.ndim() != x.ndim() || b.shape(0) != x.shape(0) ||
                    (b.ndim() == 2 && b.shape(1) != x.shape(1)))
                    throw std::invalid_argument("x and b should have the same dimensions.");
                if (b.shape(0) != self.n()) throw std::invalid_argument("b and x should have n_rows rows.");
                const int k = b.ndim() == 2 ? (int)b.shape(1) : 1;
                if (k > 128) throw std::invalid_argument("The number of RHS should be less than 128.");
                if (!self.gpu()) throw std::invalid_argument("Refined solves only exist for GPU solvers.");
                int steps = 0;
                double residual = 0.0;
                throw_on_error(cholespy_solver_solve_refined_tol(self.gpu(), k, static_cast<const double *>(b.data()),
                                                                 static_cast<double *>(x.data()), tol, max_steps,
                                                                 nullptr, &steps, &residual));
                return py::make_tuple(steps, residual);
            },
            py::arg("b"), py::arg("x"), py::arg("tol"), py::arg("max_steps") = 10)
        .def("_plan_info",
             [](const Class &self) {
                 const std::vector<int64_t> v = self.plan_info();
                 const char *names[] = {"blocks", "m_entries", "levels_fwd", "levels_bwd", "fwd_tasks",
                                        "bwd_tasks", "slots", "max_rows", "fwd_tasks_wide", "bwd_tasks_wide"};
                 py::dict d;
                 for (size_t i = 0; i < v.size(); ++i) d[names[i]] = v[i];
                 return d;
             })
        .def_property_readonly("_n", [](const Class &self) { return self.n(); })
        .def_property_readonly("_is_cpu", [](const Class &self) { return self.is_cpu(); });
}

template <typename Float>
void declare_batch(py::module_ &m, const char *name) {
    using Class = BatchedCholeskySolver<Float>;
    py::class_<Class>(m, name,
                      "Independent sparse SPD systems factored on the host and solved together on the GPU "
                      "(one launch per triangular pass).")
        .def(py::init([](py::list systems) {
                 std::vector<SystemInput> in;
                 for (py::handle item : systems) {
                     py::tuple t = py::reinterpret_borrow<py::tuple>(item);
                     if (t.size() != 5)
                         throw py::type_error("each system is (n_rows, ii, jj, x, type)");
                     SystemInput si;
                     si.n = t[0].cast<int>();
                     py::object ii_o = as_exportable(t[1]), jj_o = as_exportable(t[2]), x_o = as_exportable(t[3]);
                     DLView ii(ii_o), jj(jj_o), x(x_o);
                     si.type = t[4].cast<MatrixType>();
                     si.ii = to_host<int>(ii);
                     si.jj = to_host<int>(jj);
                     si.x = to_host<double>(x);
                     in.push_back(std::move(si));
                 }
                 py::gil_scoped_release release;
                 return std::make_unique<Class>(in);
             }),
             py::arg("systems"))
        .def(
            "solve",
            [](Class &self, std::vector<int> n_rhs, py::object b_o, py::object x_o, uintptr_t stream) {
                DLView b(b_o), x(x_o);
                const int bits = 8 * (int)sizeof(Float);
                for (const DLView *v : {&b, &x})
                    if (v->device() != Dev::GPU || !v->c_contig() || v->dtype().code != dl::kFloat ||
                        v->dtype().bits != bits)
                        throw py::type_error("solve(): b and x must be packed C-contiguous GPU arrays of the "
                                             "batch dtype.");
                if ((int)n_rhs.size() != self.count()) throw std::invalid_argument("n_rhs needs one entry per system.");
                int64_t need = 0;
                for (int i = 0; i < self.count(); ++i) need += (int64_t)self.sizes()[i] * n_rhs[i];
                if (b.size() != need || x.size() != need)
                    throw std::invalid_argument("b and x must hold sum(n_i * n_rhs_i) elements.");
                self.solve(n_rhs, b.data(), x.data(), reinterpret_cast<void *>(stream));
            },
            py::arg("n_rhs"), py::arg("b"), py::arg("x"), py::arg("stream") = 0)
        .def("_factor",
             [](const Class &self, int i) {
                 const Factor &f = self.factor(i);
                 py::dict d;
                 d["perm"] = py::array_t<int>(f.perm.size(), f.perm.data());
                 d["colptr"] = py::array_t<int>(f.colptr.size(), f.colptr.data());
                 d["rowidx"] = py::array_t<int>(f.rowidx.size(), f.rowidx.data());
                 d["val"] = py::array_t<double>(f.val.size(), f.val.data());
                 return d;
             })
        .def("_check", [](const Class &self) { throw_on_error(cholespy_solver_check(self.gpu())); })
        .def("_plan_info",
             [](const Class &self) {
                 std::vector<int64_t> v(10, 0);
                 throw_on_error(cholespy_solver_plan_info(self.gpu(), v.data(), (int)v.size()));
                 const char *names[] = {"blocks", "m_entries", "levels_fwd", "levels_bwd", "fwd_tasks",
                                        "bwd_tasks", "slots", "max_rows", "fwd_tasks_wide", "bwd_tasks_wide"};
                 py::dict d;
                 for (size_t i = 0; i < v.size(); ++i) d[names[i]] = v[i];
                 return d;
             })
        .def("_plan_arrays", [](const Class &self) { return plan_arrays(self.gpu()); })
        .def("_set_trace", [](Class &self, bool on) { throw_on_error(cholespy_solver_set_trace(self.gpu(), on ? 1 : 0)); },
             py::arg("on"))
        .def("_trace",
             [](Class &self, bool forward) {
                 std::vector<int64_t> info(10, 0);
                 throw_on_error(cholespy_solver_plan_info(self.gpu(), info.data(), (int)info.size()));
                 const int64_t words = 4 * (forward ? std::max(info[4], info[8]) : std::max(info[5], info[9]));
                 py::array_t<uint64_t> out(words);
                 throw_on_error(cholespy_solver_get_trace(self.gpu(), forward ? 1 : 0, out.mutable_data(), words));
                 return out;
             },
             py::arg("forward"))
        .def_property_readonly("sizes", [](const Class &self) { return self.sizes(); });
}

}  // namespace

// dL/d(stored entries) on the GPU (grad.hip through the C ABI): rows / cols
// int32, lam / x row-major float arrays of one dtype, grad float32/64.
void entry_gradient(py::object rows_o, py::object cols_o, py::object lam_o, py::object x_o,
                    py::object grad_o) {
    DLView rows(rows_o), cols(cols_o), lam(lam_o), x(x_o), grad(grad_o);
    for (const DLView *v : {&rows, &cols, &lam, &x, &grad})
        if (v->device() != Dev::GPU || !v->c_contig())
            throw py::type_error("_entry_gradient(): all arrays must be C-contiguous GPU arrays.");
    if (rows.dtype().code != dl::kInt || rows.dtype().bits != 32 || cols.dtype().code != dl::kInt ||
        cols.dtype().bits != 32)
        throw py::type_error("_entry_gradient(): rows and cols must be int32.");
    if (lam.dtype().code != dl::kFloat || x.dtype().code != dl::kFloat || grad.dtype().code != dl::kFloat ||
        lam.dtype().bits != x.dtype().bits)
        throw py::type_error("_entry_gradient(): lam, x and grad must be float arrays (lam, x same dtype).");
    if (rows.size() != cols.size() || rows.size() != grad.size() || lam.ndim() != x.ndim() ||
        lam.size() != x.size())
        throw std::invalid_argument("_entry_gradient(): inconsistent sizes.");
    const int k = x.ndim() == 2 ? (int)x.shape(1) : 1;
    throw_on_error(cholespy_entry_gradient((int)x.shape(0), rows.size(), static_cast<const int *>(rows.data()),
                                           static_cast<const int *>(cols.data()), k, lam.data(), x.data(),
                                           lam.dtype().bits / 8, grad.data(), grad.dtype().bits / 8,
                                           nullptr));
}

PYBIND11_MODULE(_cholespy_core, m) {
    m.def("_probe_handoff", []() {
        double r[6] = {0};
        throw_on_error(cholespy_probe_handoff(r, 6));
        py::dict d;
        d["handoff_relaxed_us"] = r[0];
        d["handoff_fenced_us"] = r[1];
        d["agent_load_ns"] = r[2];
        d["barrier_ns"] = r[3];
        d["ticket_atomic_ns"] = r[4];
        d["launch_us"] = r[5];
        return d;
    }, "Microbenchmark of the hand-off costs the block schedule is built around (cholespy_probe_handoff).");
    m.def("_entry_gradient", &entry_gradient, py::arg("rows"), py::arg("cols"), py::arg("lam"),
          py::arg("x"), py::arg("grad"));
    py::enum_<MatrixType>(m, "MatrixType", doc_matrix_type)
        .value("CSC", MatrixType::CSC)
        .value("CSR", MatrixType::CSR)
        .value("COO", MatrixType::COO);

    declare_solver<float>(m, "CholeskySolverF", doc_cholesky_f);
    declare_solver<double>(m, "CholeskySolverD", doc_cholesky_d);
    declare_batch<float>(m, "BatchedSolverF");
    declare_batch<double>(m, "BatchedSolverD");

    // Release library state when the module is unloaded (src/main.cpp:126-129).
    m.add_object("_shutdown_guard", py::capsule(reinterpret_cast<void *>(1), [](void *) {
                     cholespy_shutdown();
                 }));
    m.attr("__version__") = "2.2.2";
    m.attr("_PATH_ROWS") = (int)CHOLESPY_PATH_ROWS;
    m.attr("_PATH_BLOCKS") = (int)CHOLESPY_PATH_BLOCKS;
}
