// CuArray<T>: an m x n row-major host array with a lazily allocated device
// mirror — the buffer type of the netcalc API (reference:
// cuarray/include/cuarray.h:24-489, cuarray/src/cuarray.cu:8-454).
//
// Same observable behaviour as the reference: zero-filled init, shallow row
// views that alias the parent (owner() == 0), deep copies (owner() == 1),
// toDevice()/toHost() returning 1 unless both host and device are allocated,
// set() bounds-checked and get() not, sort()/argsort() descending, .npy
// load/save. The device side goes through libnsx (nsx_alloc / nsx_free /
// pinned-staged nsx_memcpy_*), not the CUDA runtime.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../../include/netsci_hip.h"

typedef int CuArrayError;

namespace npy {

template <class T>
const char *descr();
template <>
inline const char *descr<float>() { return "<f4"; }
template <>
inline const char *descr<double>() { return "<f8"; }
template <>
inline const char *descr<int>() { return "<i4"; }

// Write a version-1.0 .npy file holding an m x n C-order array.
template <class T>
void save(const std::string &fname, const T *data, long long m, long long n) {
    std::ostringstream hdr;
    hdr << "{'descr': '" << descr<T>() << "', 'fortran_order': False, 'shape': (" << m
        << ", " << n << "), }";
    std::string h = hdr.str();
    size_t total = 10 + h.size() + 1;
    size_t pad = (64 - total % 64) % 64;
    h.append(pad, ' ');
    h.push_back('\n');
    std::ofstream f(fname, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + fname + " for writing");
    const unsigned char magic[8] = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0};
    f.write((const char *)magic, 8);
    uint16_t hl = (uint16_t)h.size();
    f.write((const char *)&hl, 2);
    f.write(h.data(), (std::streamsize)h.size());
    f.write((const char *)data, (std::streamsize)(sizeof(T) * m * n));
}

// Read a .npy file (1-D or 2-D, C order, little-endian f4/f8/i4/i8) and
// convert its values to T. shape gets {m, n} (1-D arrays give {1, len}).
template <class T>
std::vector<T> load(const std::string &fname, long long &m, long long &n) {
    std::ifstream f(fname, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open " + fname);
    char magic[6];
    f.read(magic, 6);
    if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0)
        throw std::runtime_error(fname + " is not a .npy file");
    unsigned char ver[2];
    f.read((char *)ver, 2);
    uint32_t hlen = 0;
    if (ver[0] == 1) {
        uint16_t h16;
        f.read((char *)&h16, 2);
        hlen = h16;
    } else {
        f.read((char *)&hlen, 4);
    }
    std::string h(hlen, '\0');
    f.read(&h[0], hlen);
    auto field = [&](const std::string &key) {
        size_t p = h.find("'" + key + "'");
        if (p == std::string::npos) throw std::runtime_error(fname + ": no " + key);
        return p + key.size() + 2;
    };
    size_t p = h.find('\'', h.find(':', field("descr")) + 1);
    std::string dt = h.substr(p + 1, h.find('\'', p + 1) - p - 1);
    if (h.find("True", field("fortran_order")) < h.find(',', field("fortran_order")))
        throw std::runtime_error(fname + ": Fortran-order arrays are not supported");
    size_t s0 = h.find('(', field("shape")), s1 = h.find(')', s0);
    std::string shp = h.substr(s0 + 1, s1 - s0 - 1);
    std::vector<long long> dims;
    std::stringstream ss(shp);
    std::string tok;
    while (std::getline(ss, tok, ',')) {
        tok.erase(std::remove(tok.begin(), tok.end(), ' '), tok.end());
        if (!tok.empty()) dims.push_back(std::stoll(tok));
    }
    if (dims.size() == 1) {
        m = 1;
        n = dims[0];
    } else if (dims.size() == 2) {
        m = dims[0];
        n = dims[1];
    } else {
        throw std::runtime_error(fname + ": only 1-D and 2-D arrays are supported");
    }
    const long long count = m * n;
    std::vector<T> out((size_t)count);
    auto read_as = [&](auto tag) {
        using S = decltype(tag);
        std::vector<S> raw((size_t)count);
        f.read((char *)raw.data(), (std::streamsize)(sizeof(S) * count));
        if (!f) throw std::runtime_error(fname + ": truncated data");
        for (long long i = 0; i < count; ++i) out[i] = (T)raw[i];
    };
    if (dt == "<f4") read_as(float());
    else if (dt == "<f8") read_as(double());
    else if (dt == "<i4") read_as(int32_t());
    else if (dt == "<i8") read_as(int64_t());
    else throw std::runtime_error(fname + ": unsupported dtype " + dt);
    return out;
}

}  // namespace npy

template <typename T>
class CuArray {
public:
    CuArray() = default;
    CuArray(const CuArray &) = delete;
    CuArray &operator=(const CuArray &) = delete;

    ~CuArray() {
        if (allocatedHost_ == 1 && owner_) deallocateHost();
        if (allocatedDevice_ == 1) deallocateDevice();
    }

    // cuarray.cu:16-42 — (re)allocate a zero-filled m x n host array.
    CuArrayError init(int m, int n) {
        if (allocatedHost_ == 1 && owner_) deallocateHost();
        if (allocatedDevice_ == 1) deallocateDevice();
        setShape(m, n);
        owner_ = 1;
        host_ = new T[(size_t)size_]();
        allocatedHost_ = 1;
        device_ = nullptr;
        return 0;
    }

    // cuarray.cu:44-63 — copy m*n values from host.
    CuArrayError init(const T *host, int m, int n) {
        init(m, n);
        std::copy(host, host + (size_t)size_, host_);
        return 0;
    }

    // cuarray.cu:314-335 — alias rows start..end of cuArray as m x n.
    CuArrayError fromCuArrayShallowCopy(CuArray<T> *cuArray, int start, int end, int m, int n) {
        if ((long long)((end - start) + 1) * cuArray->n_ != (long long)m * n) return 1;
        if (allocatedHost_ == 1 && owner_) deallocateHost();
        if (allocatedDevice_ == 1) deallocateDevice();
        owner_ = 0;
        setShape(m, n);
        allocatedHost_ = 1;
        allocatedDevice_ = 0;
        host_ = cuArray->host_ + (size_t)start * cuArray->n_;
        return 0;
    }

    // cuarray.cu:337-363 — copy rows start..end of cuArray as m x n.
    CuArrayError fromCuArrayDeepCopy(CuArray<T> *cuArray, int start, int end, int m, int n) {
        if ((long long)((end - start) + 1) * cuArray->n_ != (long long)m * n) return 1;
        std::vector<T> tmp(cuArray->host_ + (size_t)start * cuArray->n_,
                           cuArray->host_ + (size_t)(end + 1) * cuArray->n_);
        init(m, n);
        std::copy(tmp.begin(), tmp.end(), host_);
        return 0;
    }

    int n() const { return n_; }
    int m() const { return m_; }
    int size() const { return size_; }
    size_t bytes() const { return bytes_; }
    T *&host() { return host_; }
    T *&device() { return device_; }
    int owner() const { return owner_; }

    CuArrayError allocateHost() {
        if (allocatedHost_ == 0 && owner_) {
            host_ = new T[(size_t)size_]();
            allocatedHost_ = 1;
        }
        return (allocatedHost_ == 1) ? 0 : 1;
    }

    CuArrayError allocateDevice() {
        if (allocatedDevice_ == 0) {
            void *p = nullptr;
            if (nsx_alloc(&p, bytes_) != 0)
                throw std::runtime_error(std::string("CuArray::allocateDevice: ") + nsx_last_error());
            device_ = (T *)p;
            allocatedDevice_ = 1;
        }
        return (allocatedDevice_ == 1) ? 0 : 1;
    }

    CuArrayError allocatedHost() const { return allocatedHost_; }
    CuArrayError allocatedDevice() const { return allocatedDevice_; }

    // cuarray.cu:124-134 — copy only when both sides exist, else return 1.
    CuArrayError toDevice() {
        if (allocatedDevice_ == 1 && allocatedHost_ == 1) {
            if (nsx_memcpy_h2d(device_, host_, bytes_) != 0)
                throw std::runtime_error(std::string("CuArray::toDevice: ") + nsx_last_error());
            return 0;
        }
        return 1;
    }

    CuArrayError toHost() {
        if (allocatedDevice_ == 1 && allocatedHost_ == 1) {
            if (nsx_memcpy_d2h(host_, device_, bytes_) != 0)
                throw std::runtime_error(std::string("CuArray::toHost: ") + nsx_last_error());
            return 0;
        }
        return 1;
    }

    CuArrayError deallocateHost() {
        if (allocatedHost_ == 1 && owner_) {
            delete[] host_;
            host_ = nullptr;
            allocatedHost_ = 0;
        }
        return (allocatedHost_ == 0) ? 0 : 1;
    }

    CuArrayError deallocateDevice() {
        if (allocatedDevice_ == 1) {
            nsx_free(device_);
            device_ = nullptr;
            allocatedDevice_ = 0;
        }
        return (allocatedDevice_ == 0) ? 0 : 1;
    }

    // cuarray.cu:184-228.
    CuArrayError fromNumpy(const T *data, int dim1) { return init(data, 1, dim1); }
    CuArrayError fromNumpy(const T *data, int dim1, int dim2) { return init(data, dim1, dim2); }

    // cuarray.cu:230-249 — get() unchecked, set() bounds-checked.
    T get(int i, int j) const { return host_[(size_t)i * n_ + j]; }
    CuArrayError set(T value, int i, int j) {
        if (i >= 0 && j >= 0 && i < m_ && j < n_) {
            host_[(size_t)i * n_ + j] = value;
            return 0;
        }
        return 1;
    }
    T &operator[](int i) const { return host_[i]; }

    // cuarray.cu:251-282.
    CuArrayError load(const std::string &fname) {
        long long m = 0, n = 0;
        std::vector<T> v = npy::load<T>(fname, m, n);
        return init(v.data(), (int)m, (int)n);
    }
    void save(const std::string &fname) { npy::save<T>(fname, host_, m_, n_); }

    // cuarray.cu:370-388 — row i sorted descending, as a new 1 x n array.
    CuArray<T> *sort(int i) {
        auto *out = new CuArray<T>();
        out->fromCuArrayDeepCopy(this, i, i, 1, n_);
        std::sort(out->host_, out->host_ + out->n_, std::greater<T>());
        return out;
    }

    // cuarray.cu:390-411 — indices of row i by descending value.
    CuArray<int> *argsort(int i) {
        auto *out = new CuArray<int>();
        out->init(1, n_);
        std::iota(out->host(), out->host() + out->n(), 0);
        std::sort(out->host(), out->host() + out->n(),
                  [this, i](int a, int b) { return get(i, a) > get(i, b); });
        return out;
    }

private:
    void setShape(int m, int n) {
        m_ = m;
        n_ = n;
        size_ = m * n;
        bytes_ = (size_t)m * (size_t)n * sizeof(T);
    }

    int m_ = 0, n_ = 0, size_ = 0;
    size_t bytes_ = 0;
    int owner_ = 1;
    int allocatedHost_ = 0, allocatedDevice_ = 0;
    T *host_ = nullptr;
    T *device_ = nullptr;
};

// cuarray.h:473-489 — a row view.
template <typename T>
class CuArrayRow {
public:
    CuArrayRow(CuArray<T> *a, int i) : data_(a->host() + (size_t)i * a->n()), n_(a->n()) {}
    T &operator[](int i) const { return data_[i]; }
    int n() const { return n_; }
    T *data() const { return data_; }

private:
    T *data_;
    int n_;
};
