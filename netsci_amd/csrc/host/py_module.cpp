// Python bindings: the `cuarray`, `netcalc` and `netchem` surfaces of the
// reference's SWIG modules (cuarray/swig/python_cuarray*.i,
// netcalc/swig/python_netcalc*.i, netchem/swig/python_netchem*.i) with the
// same names, keyword names and numpy conventions. Exceptions map to
// Python: std::invalid_argument -> ValueError, std::out_of_range -> IndexError,
// std::runtime_error -> RuntimeError.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <sstream>
#include <string>

#include "cuarray.h"
#include "netcalc.h"
#include "netchem.h"

namespace py = pybind11;
using namespace pybind11::literals;

namespace {

// python_cuarray_typemaps.i:106-141 — rows of values, floats cut to four
// decimals, padded to the widest integer part.
template <class T>
std::string repr_values(const T *data, int m, int n) {
    if (m * n == 0) return "null";
    const bool fp = std::to_string(data[0]).find('.') != std::string::npos;
    const size_t width =
        std::to_string(static_cast<int>(*std::max_element(data, data + (size_t)m * n))).size();
    std::string out;
    for (int i = 0; i < m; ++i) {
        for (int j = 0; j < n; ++j) {
            const T v = data[(size_t)i * n + j];
            std::string s;
            size_t ilen;
            if (fp) {
                std::string tmp = std::to_string(v);
                s = tmp.substr(0, tmp.find('.') + 5);
                ilen = std::to_string(static_cast<int>(v)).size();
            } else {
                s = std::to_string(v);
                ilen = s.size();
            }
            out += s;
            if (width > ilen) out += std::string(width - ilen, ' ');
            if (j < n - 1) out += " ";
        }
        if (i < m - 1) out += "\n";
    }
    return out;
}

template <class T>
void bind_cuarray(py::module_ &m, const char *name, const char *row_name) {
    using A = CuArray<T>;
    using Row = CuArrayRow<T>;
    py::class_<Row>(m, row_name)
        .def("__getitem__",
             [](Row &r, int i) {
                 if (i < 0 || i >= r.n() || r.n() == 0)
                     throw std::out_of_range("CuArrayRow index out of range");
                 return r[i];
             })
        .def("__setitem__",
             [](Row &r, int i, T v) {
                 if (i < 0 || i >= r.n() || r.n() == 0)
                     throw std::out_of_range("CuArrayRow index out of range");
                 r[i] = v;
             })
        .def("__len__", &Row::n)
        .def("__repr__", [](Row &r) { return repr_values<T>(r.data(), 1, r.n()); });

    py::class_<A>(m, name)
        .def(py::init<>())
        .def("init", [](A &a, int m, int n) { return a.init(m, n); }, "m"_a, "n"_a)
        .def("fromCuArray", &A::fromCuArrayDeepCopy, "cuArray"_a, "start"_a, "end"_a, "m"_a,
             "n"_a)
        .def("m", &A::m)
        .def("n", &A::n)
        .def("size", &A::size)
        .def("bytes", &A::bytes)
        .def("owner", &A::owner)
        .def("allocateHost", &A::allocateHost)
        .def("allocateDevice", &A::allocateDevice)
        .def("allocatedHost", &A::allocatedHost)
        .def("allocatedDevice", &A::allocatedDevice)
        .def("deallocateHost", &A::deallocateHost)
        .def("deallocateDevice", &A::deallocateDevice)
        .def("toDevice", &A::toDevice)
        .def("toHost", &A::toHost)
        .def("get", &A::get, "i"_a, "j"_a)
        .def("set", &A::set, "value"_a, "i"_a, "j"_a)
        .def("load", &A::load, "fname"_a)
        .def("save", &A::save, "fname"_a)
        .def("sort", [](A &a, int i) { return a.sort(i); }, "i"_a,
             py::return_value_policy::take_ownership)
        .def("argsort", [](A &a, int i) { return a.argsort(i); }, "i"_a,
             py::return_value_policy::take_ownership)
        .def(
            "fromNumpy1D",
            [](A &a, py::array_t<T, py::array::c_style | py::array::forcecast> arr) {
                if (arr.ndim() != 1) throw std::invalid_argument("expected a 1-D array");
                return a.fromNumpy(arr.data(), (int)arr.shape(0));
            },
            "NUMPY_ARRAY"_a)
        .def(
            "fromNumpy2D",
            [](A &a, py::array_t<T, py::array::c_style | py::array::forcecast> arr) {
                if (arr.ndim() != 2) throw std::invalid_argument("expected a 2-D array");
                return a.fromNumpy(arr.data(), (int)arr.shape(0), (int)arr.shape(1));
            },
            "NUMPY_ARRAY"_a)
        .def("toNumpy1D",
             [](A &a) {
                 py::array_t<T> out(a.size());
                 if (a.size()) std::copy(a.host(), a.host() + a.size(), out.mutable_data());
                 return out;
             })
        .def("toNumpy2D",
             [](A &a) {
                 py::array_t<T> out({a.m(), a.n()});
                 if (a.size()) std::copy(a.host(), a.host() + a.size(), out.mutable_data());
                 return out;
             })
        .def(
            "__getitem__",
            [](A &a, int i) {
                if (!a.allocatedHost()) throw std::runtime_error("CuArray memory not allocated");
                if (i < 0 || i >= a.m()) throw std::out_of_range("CuArray index out of range");
                return Row(&a, i);
            },
            py::keep_alive<0, 1>())
        .def("__setitem__",
             [](A &a, int i, Row &r) {
                 if (!a.allocatedHost()) throw std::runtime_error("CuArray memory not allocated");
                 if (i < 0 || i >= a.m()) throw std::out_of_range("CuArray index out of range");
                 for (int j = 0; j < r.n(); ++j) a.set(r[j], i, j);
             })
        .def("__len__", &A::m)
        .def("__repr__", [](A &a) {
            if (!a.allocatedHost()) return std::string("null");
            return repr_values<T>(a.host(), a.m(), a.n());
        });
}

// netchem/include/{atom,atoms,node,network}.h as wrapped by
// netchem/swig/python_netchem*.i: accessor methods, Atoms iterable.
void bind_netchem(py::module_ &m) {
    using namespace netchem;
    const auto ref = py::return_value_policy::reference_internal;
    py::class_<Atom>(m, "Atom")
        .def("index", [](const Atom &a) { return a.index; })
        .def("serial", [](const Atom &a) { return a.serial; })
        .def("name", [](const Atom &a) { return a.name; })
        .def("element", [](const Atom &a) { return a.element; })
        .def("residueName", [](const Atom &a) { return a.residueName; })
        .def("residueId", [](const Atom &a) { return a.residueId; })
        .def("chainId", [](const Atom &a) { return a.chainId; })
        .def("segmentId", [](const Atom &a) { return a.segmentId; })
        .def("occupancy", [](const Atom &a) { return a.occupancy; })
        .def("temperatureFactor", [](const Atom &a) { return a.temperatureFactor; })
        .def("tag", [](const Atom &a) { return a.tag; })
        .def("mass", [](const Atom &a) { return a.mass; })
        .def("__repr__", [](const Atom &a) {
            return "Atom(" + std::to_string(a.index) + ", " + a.name + ", " + a.tag + ")";
        });
    py::class_<Atoms>(m, "Atoms")
        .def("numAtoms", &Atoms::numAtoms)
        .def("numUniqueTags", &Atoms::numUniqueTags)
        .def("atoms", &Atoms::atoms, ref)
        .def("at", &Atoms::at, "atomIndex"_a, ref)
        .def("__len__", &Atoms::numAtoms)
        .def("__getitem__", &Atoms::at, ref)
        .def("__iter__", [](Atoms &a) { return py::iter(py::cast(a.atoms(), py::return_value_policy::reference)); },
             py::keep_alive<0, 1>());
    py::class_<Node>(m, "Node")
        .def("index", [](const Node &n) { return n.index; })
        .def("tag", [](const Node &n) { return n.tag; })
        .def("numAtoms", &Node::numAtoms)
        .def("totalMass", [](const Node &n) { return n.totalMass; })
        .def("atoms", [](const Node &n) { return n.atoms; }, ref)
        .def("__repr__", [](const Node &n) {
            return "Node(" + std::to_string(n.index) + ", " + n.tag + ")";
        });
    py::class_<Network>(m, "Network")
        .def(py::init<>())
        .def("init", &Network::init, "trajectoryFile"_a, "topologyFile"_a, "firstFrame"_a,
             "lastFrame"_a, "stride"_a = 1)
        .def("numNodes", &Network::numNodes)
        .def("numFrames", &Network::numFrames)
        .def("nodeCoordinates", py::overload_cast<>(&Network::nodeCoordinates), ref)
        .def("nodeCoordinates", py::overload_cast<const std::string &>(&Network::nodeCoordinates),
             "nodeCoordinatesFile"_a)
        .def("nodes", &Network::nodes, ref)
        .def("atoms", &Network::atoms, ref)
        .def("nodeFromAtomIndex", &Network::nodeFromAtomIndex, "atomIndex"_a, ref);
    m.def("dcdFrameCount", [](const std::string &path) { return read_dcd_layout(path).nframes; },
          "trajectoryFile"_a);
    m.def("dcdAtomCount", [](const std::string &path) { return read_dcd_layout(path).natoms; },
          "trajectoryFile"_a);
    m.def("atomicWeight", &atomic_weight, "symbol"_a);
}

}  // namespace

PYBIND11_MODULE(_netsci, m) {
    m.doc() = "MI355X engine behind the netcalc / cuarray API";
    bind_cuarray<float>(m, "FloatCuArray", "FloatCuArrayRow");
    bind_cuarray<int>(m, "IntCuArray", "IntCuArrayRow");
    bind_netchem(m);

    using namespace netcalc;
    m.attr("GPU_PLATFORM") = GPU_PLATFORM;
    m.attr("CPU_PLATFORM") = CPU_PLATFORM;

    m.def("mutualInformation",
          py::overload_cast<CuArray<float> *, CuArray<float> *, CuArray<int> *, int, int, int, int,
                            int>(&mutualInformation),
          "X"_a, "I"_a, "ab"_a, "k"_a, "n"_a, "xd"_a, "d"_a, "platform"_a);
    m.def("mutualInformationWithCheckpointing",
          py::overload_cast<CuArray<float> *, CuArray<float> *, CuArray<int> *, int, int, int, int,
                            int, int, std::string>(&mutualInformation),
          "X"_a, "I"_a, "ab"_a, "k"_a, "n"_a, "xd"_a, "d"_a, "platform"_a,
          "checkpointFrequency"_a, "checkpointFileName"_a);
    m.def("mutualInformationGpu", &mutualInformationGpu, "Xa"_a, "Xb"_a, "k"_a, "n"_a, "xd"_a,
          "d"_a);
    m.def("mutualInformationCpu", &mutualInformationCpu, "Xa"_a, "Xb"_a, "k"_a, "n"_a, "xd"_a,
          "d"_a);
    m.def("generalizedCorrelation",
          py::overload_cast<CuArray<float> *, CuArray<float> *, CuArray<int> *, int, int, int, int,
                            int>(&generalizedCorrelation),
          "X"_a, "R"_a, "ab"_a, "k"_a, "n"_a, "xd"_a, "d"_a, "platform"_a);
    m.def("generalizedCorrelationWithCheckpointing",
          py::overload_cast<CuArray<float> *, CuArray<float> *, CuArray<int> *, int, int, int, int,
                            int, int, std::string>(&generalizedCorrelation),
          "X"_a, "R"_a, "ab"_a, "k"_a, "n"_a, "xd"_a, "d"_a, "platform"_a,
          "checkpointFrequency"_a, "checkpointFileName"_a);
    m.def("generalizedCorrelationGpu", &generalizedCorrelationGpu, "Xa"_a, "Xb"_a, "k"_a, "n"_a,
          "xd"_a, "d"_a);
    m.def("generalizedCorrelationCpu", &generalizedCorrelationCpu, "Xa"_a, "Xb"_a, "k"_a, "n"_a,
          "xd"_a, "d"_a);
    m.def("generateRestartAbFromCheckpointFile", &generateRestartAbFromCheckpointFile, "ab"_a,
          "restartAb"_a, "checkpointFileName"_a);
    m.def("hedetniemiAllShortestPaths", &hedetniemiAllShortestPaths, "A"_a, "H"_a, "paths"_a,
          "maxPathLength"_a, "tolerance"_a, "platform"_a);
    m.def("hedetniemiAllShortestPathsGpu", &hedetniemiAllShortestPathsGpu, "A"_a, "H"_a,
          "paths"_a, "tolerance"_a, "maxPathLength"_a);
    m.def("hedetniemiAllShortestPathsCpu", &hedetniemiAllShortestPathsCpu, "A"_a, "H"_a,
          "paths"_a, "tolerance"_a, "maxPathLength"_a);
    m.def("hedetniemiAllShortestPathLengths", &hedetniemiAllShortestPathLengths, "A"_a, "H"_a,
          "maxPathLength"_a, "platform"_a);
    m.def("hedetniemiAllShortestPathLengthsGpu", &hedetniemiAllShortestPathLengthsGpu, "A"_a,
          "H"_a, "maxPathLength"_a);
    m.def("correlationToAdjacency", &correlationToAdjacency, "A"_a, "C"_a, "n"_a, "platform"_a);
    m.def("correlationToAdjacencyGpu", &correlationToAdjacencyGpu, "A"_a, "C"_a, "n"_a);
    m.def("mean", &mean, "a"_a, "u"_a, "m"_a, "n"_a, "platform"_a);
    m.def("meanGpu", &meanGpu, "a"_a, "u"_a, "m"_a, "n"_a);
    m.def("standardDeviation", &standardDeviation, "a"_a, "u"_a, "sigma"_a, "m"_a, "n"_a,
          "platform"_a);
    m.def("standardDeviationGpu", &standardDeviationGpu, "a"_a, "u"_a, "sigma"_a, "m"_a, "n"_a);
    m.def(
        "recoverSingleShortestPath",
        [](CuArray<int> *paths, int maxPathLength, int i, int j) {
            std::vector<int> p = recoverSingleShortestPath(paths, maxPathLength, i, j);
            py::array_t<int> out((py::ssize_t)p.size());
            std::copy(p.begin(), p.end(), out.mutable_data());
            return out;
        },
        "paths"_a, "maxPathLength"_a, "i"_a, "j"_a);
    m.def("windowedCorrelation", &windowedCorrelation, "X"_a, "R"_a, "ab"_a, "windows"_a, "k"_a,
          "nTotal"_a, "d"_a, "gc"_a = 1);
}
