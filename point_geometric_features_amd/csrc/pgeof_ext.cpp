// Python module `pgeof_ext`: the reference's binding surface
// (src/pgeof_ext.cpp:12-178, NB_MODULE(pgeof_ext, m)) over the C ABI of
// libpgeof_hip.so.  Same function names, keyword names, defaults, dtype rules
// (float32 / uint32 without implicit conversion -> TypeError) and ValueError
// conditions.  Written with pybind11 because nanobind is not available to
// this build; the GIL is released around every device call.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pgeof_hip.h"

namespace py = pybind11;
using namespace py::literals;

namespace
{

// EFeatureID, include/pca.hpp:47-63
enum EFeatureID
{
    Linearity = 0,
    Planarity,
    Scattering,
    VerticalityPGEOF,
    Normal_x,
    Normal_y,
    Normal_z,
    Length,
    Surface,
    Volume,
    Curvature,
    K_optimal,
    Verticality,
    Eigentropy
};

[[noreturn]] void raise(int rc)
{
    const std::string msg = pgeof_last_error();
    if (rc == PGEOF_EINVAL) throw py::value_error(msg);
    if (rc == PGEOF_ENOMEM) throw std::bad_alloc();
    throw std::runtime_error(msg);
}

void check(int rc)
{
    if (rc != PGEOF_OK) raise(rc);
}

template <typename T>
const char* dtype_name()
{
    if (std::is_same<T, float>::value) return "float32";
    if (std::is_same<T, double>::value) return "float64";
    return "uint32";
}

// Accept a numpy array of exactly dtype T (no conversion, like nanobind's
// .noconvert()), return a C-contiguous view or copy.
template <typename T>
py::array_t<T, py::array::c_style> exact(const py::handle& obj, const char* name, int ndim, int last_dim = -1)
{
    if (!py::isinstance<py::array>(obj))
        throw py::type_error(std::string("incompatible function arguments: '") + name +
                             "' must be a numpy.ndarray of " + dtype_name<T>());
    py::array arr = py::reinterpret_borrow<py::array>(obj);
    if (!arr.dtype().is(py::dtype::of<T>()))
        throw py::type_error(std::string("incompatible function arguments: '") + name + "' must have dtype " +
                             dtype_name<T>() + ", got " + std::string(py::str(arr.dtype())));
    if (arr.ndim() != ndim)
        throw py::type_error(std::string("incompatible function arguments: '") + name + "' must be " +
                             std::to_string(ndim) + "-dimensional");
    if (last_dim >= 0 && arr.shape(ndim - 1) != last_dim)
        throw py::type_error(std::string("incompatible function arguments: '") + name + "' must have shape (n, " +
                             std::to_string(last_dim) + ")");
    return py::array_t<T, py::array::c_style>::ensure(arr);
}

void progress_done(bool verbose)
{
    // The reference prints a progress line every 10000 points and a final
    // newline (include/pgeof.hpp:32-48); the device path reports completion.
    if (verbose)
    {
        py::print("100% done");
    }
}

py::array_t<float> compute_features(py::handle xyz_o, py::handle nn_o, py::handle ptr_o, py::int_ k_min_o,
                                    bool verbose)
{
    auto xyz = exact<float>(xyz_o, "xyz", 2, 3);
    auto nn = exact<uint32_t>(nn_o, "nn", 1);
    auto ptr = exact<uint32_t>(ptr_o, "nn_ptr", 1);
    const long long k_min_s = k_min_o.cast<long long>();
    if (k_min_s < 0) throw py::type_error("incompatible function arguments: 'k_min' must be a non-negative int");
    const uint64_t k_min = (uint64_t)k_min_s;
    if (ptr.shape(0) < 1) throw py::value_error("nn_ptr must hold at least one entry");
    const ssize_t n = ptr.shape(0) - 1;
    py::array_t<float> out({n, (ssize_t)11});
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = pgeof_compute_features(xyz.data(), xyz.shape(0), nn.data(), nn.shape(0), ptr.data(), ptr.shape(0), k_min,
                                    out.mutable_data());
    }
    check(rc);
    progress_done(verbose);
    return out;
}

py::array_t<float> compute_features_multiscale(py::handle xyz_o, py::handle nn_o, py::handle ptr_o,
                                               std::vector<uint32_t> k_scales, bool verbose)
{
    auto xyz = exact<float>(xyz_o, "xyz", 2, 3);
    auto nn = exact<uint32_t>(nn_o, "nn", 1);
    auto ptr = exact<uint32_t>(ptr_o, "nn_ptr", 1);
    uint32_t prev = 1;
    for (uint32_t s : k_scales)
    {
        if (s < prev) throw py::value_error("k_scales should be > 1 and sorted in ascending order");
        prev = s;
    }
    if (ptr.shape(0) < 1) throw py::value_error("nn_ptr must hold at least one entry");
    const ssize_t n = ptr.shape(0) - 1;
    py::array_t<float> out({n, (ssize_t)k_scales.size(), (ssize_t)11});
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = pgeof_compute_features_multiscale(xyz.data(), xyz.shape(0), nn.data(), nn.shape(0), ptr.data(),
                                               ptr.shape(0), k_scales.data(), (int64_t)k_scales.size(),
                                               out.mutable_data());
    }
    check(rc);
    progress_done(verbose);
    return out;
}

py::array_t<float> compute_features_optimal(py::handle xyz_o, py::handle nn_o, py::handle ptr_o, uint32_t k_min,
                                            uint32_t k_step, uint32_t k_min_search, bool verbose)
{
    auto xyz = exact<float>(xyz_o, "xyz", 2, 3);
    auto nn = exact<uint32_t>(nn_o, "nn", 1);
    auto ptr = exact<uint32_t>(ptr_o, "nn_ptr", 1);
    if (k_min < 1 && k_min_search < 1) throw py::value_error("k_min and k_min_search should be > 1");
    if (ptr.shape(0) < 1) throw py::value_error("nn_ptr must hold at least one entry");
    const ssize_t n = ptr.shape(0) - 1;
    py::array_t<float> out({n, (ssize_t)12});
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = pgeof_compute_features_optimal(xyz.data(), xyz.shape(0), nn.data(), nn.shape(0), ptr.data(),
                                            ptr.shape(0), k_min, k_step, k_min_search, out.mutable_data());
    }
    check(rc);
    progress_done(verbose);
    return out;
}

py::tuple knn_search(py::handle data_o, py::handle query_o, uint32_t knn)
{
    auto data = exact<float>(data_o, "data", 2, 3);
    auto query = exact<float>(query_o, "query", 2, 3);
    if ((int64_t)knn > (int64_t)data.shape(0))
        throw py::value_error("knn size is greater than the data point cloud size");
    const ssize_t m = query.shape(0);
    py::array_t<uint32_t> idx({m, (ssize_t)knn});
    py::array_t<float> d2({m, (ssize_t)knn});
    // the same buffer for data and query selects the self-search path
    const float* qp = (query.data() == data.data() && query.shape(0) == data.shape(0)) ? data.data() : query.data();
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = pgeof_knn_search(data.data(), data.shape(0), qp, m, knn, idx.mutable_data(), d2.mutable_data());
    }
    check(rc);
    return py::make_tuple(idx, d2);
}

py::tuple radius_search(py::handle data_o, py::handle query_o, float search_radius, uint32_t max_knn)
{
    auto data = exact<float>(data_o, "data", 2, 3);
    auto query = exact<float>(query_o, "query", 2, 3);
    if ((int64_t)max_knn > (int64_t)data.shape(0))
        throw py::value_error("max knn size is greater than the data point cloud size");
    const ssize_t m = query.shape(0);
    py::array_t<int32_t> idx({m, (ssize_t)max_knn});
    py::array_t<float> d2({m, (ssize_t)max_knn});
    const float* qp = (query.data() == data.data() && query.shape(0) == data.shape(0)) ? data.data() : query.data();
    int rc;
    {
        py::gil_scoped_release nogil;
        rc = pgeof_radius_search(data.data(), data.shape(0), qp, m, search_radius, max_knn, idx.mutable_data(),
                                 d2.mutable_data());
    }
    check(rc);
    return py::make_tuple(idx, d2);
}

template <typename T>
py::array_t<T> compute_features_selected(py::handle xyz_o, T search_radius, uint32_t max_knn,
                                         std::vector<EFeatureID> selected)
{
    auto xyz = exact<T>(xyz_o, "xyz", 2, 3);
    const ssize_t n = xyz.shape(0);
    std::vector<int32_t> sel(selected.begin(), selected.end());
    py::array_t<T> out({n, (ssize_t)sel.size()});
    int rc;
    {
        py::gil_scoped_release nogil;
        if constexpr (std::is_same<T, double>::value)
            rc = pgeof_compute_features_selected_f64(xyz.data(), n, search_radius, max_knn, sel.data(),
                                                     (int64_t)sel.size(), out.mutable_data());
        else
            rc = pgeof_compute_features_selected_f32(xyz.data(), n, search_radius, max_knn, sel.data(),
                                                     (int64_t)sel.size(), out.mutable_data());
    }
    check(rc);
    return out;
}

// overload dispatch without implicit conversion of xyz (the reference's two
// noconvert overloads, double first: src/pgeof_ext.cpp:148-177)
py::object compute_features_selected_any(py::handle xyz_o, py::object search_radius, uint32_t max_knn,
                                         std::vector<EFeatureID> selected)
{
    if (py::isinstance<py::array>(xyz_o))
    {
        py::array a = py::reinterpret_borrow<py::array>(xyz_o);
        if (a.dtype().is(py::dtype::of<double>()))
            return compute_features_selected<double>(xyz_o, search_radius.cast<double>(), max_knn, selected);
        if (a.dtype().is(py::dtype::of<float>()))
            return compute_features_selected<float>(xyz_o, search_radius.cast<float>(), max_knn, selected);
    }
    throw py::type_error("compute_features_selected(): incompatible function arguments: 'xyz' must be a float64 or "
                         "float32 numpy.ndarray of shape (n, 3)");
}

}  // namespace

PYBIND11_MODULE(pgeof_ext, m)
{
    m.doc() = "Compute, for each point in a 3D point cloud, local geometric features on an AMD Instinct GPU (gfx950)";

    py::enum_<EFeatureID>(m, "EFeatureID")
        .value("Linearity", Linearity)
        .value("Planarity", Planarity)
        .value("Scattering", Scattering)
        .value("VerticalityPGEOF", VerticalityPGEOF)
        .value("Normal_x", Normal_x)
        .value("Normal_y", Normal_y)
        .value("Normal_z", Normal_z)
        .value("Length", Length)
        .value("Surface", Surface)
        .value("Volume", Volume)
        .value("Curvature", Curvature)
        .value("K_optimal", K_optimal)
        .value("Verticality", Verticality)
        .value("Eigentropy", Eigentropy)
        .export_values();

    m.def("compute_features", &compute_features, "xyz"_a, "nn"_a, "nn_ptr"_a, "k_min"_a = 1, "verbose"_a = false,
          R"(Compute a set of geometric features for a point cloud from a precomputed list of neighbors.

    Features (columns): linearity, planarity, scattering, verticality, normal x/y/z
    (oriented towards positive z), length, surface, volume, curvature.
    :param xyz: float32 array (n, 3).
    :param nn: uint32 1-D array, flattened neighbor indices.
    :param nn_ptr: uint32 1-D array (n_points+1); neighbors of point i are nn[nn_ptr[i]:nn_ptr[i+1]].
    :param k_min: minimum number of neighbors; points with fewer get all-zero features.
    :param verbose: print a completion line.
    :return: float32 array (n_points, 11).)");

    m.def("compute_features_multiscale", &compute_features_multiscale, "xyz"_a, "nn"_a, "nn_ptr"_a, "k_scales"_a,
          "verbose"_a = false,
          R"(Compute the geometric features at several neighborhood sizes (prefixes of each CSR row).

    :param k_scales: non-decreasing sizes, each >= 1.
    :return: float32 array (n_points, n_scales, 11); scales larger than a row's length are 0.)");

    m.def("compute_features_optimal", &compute_features_optimal, "xyz"_a, "nn"_a, "nn_ptr"_a, "k_min"_a = 1,
          "k_step"_a = 1, "k_min_search"_a = 1, "verbose"_a = false,
          R"(Compute the geometric features at the neighborhood size of minimal eigentropy (Weinmann et al., 2015).

    :return: float32 array (n_points, 12); column 11 holds the optimal neighborhood size.)");

    m.def("knn_search", &knn_search, "data"_a, "query"_a, "knn"_a,
          R"(For each query point, the knn nearest data points (squared L2 in float32, ties by index).

    :return: (indices uint32 (n_query, knn), squared distances float32 (n_query, knn)).)");

    m.def("radius_search", &radius_search, "data"_a, "query"_a, "search_radius"_a, "max_knn"_a,
          R"(For each query point, up to max_knn nearest data points with squared distance < search_radius**2.

    :return: (indices int32 (n_query, max_knn) padded with -1, squared distances float32 padded with 0).)");

    m.def("compute_features_selected", &compute_features_selected_any, "xyz"_a, "search_radius"_a, "max_knn"_a,
          "selected_features"_a,
          R"(Compute a selected set of features (EFeatureID list) on radius neighborhoods, in the dtype of xyz.)");
}
