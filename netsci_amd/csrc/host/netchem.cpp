// netchem host code: PDB atom records, DCD frames, per-residue mass-weighted
// node coordinates. See netchem.h for the reference files this follows.
#include "netchem.h"

#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <unordered_map>

namespace netchem {
namespace {

// Standard atomic weights in kg/mol, by atomic number. The values are the
// reference's (netchem/include/chemical_properties.h): a node's mass is
// (float)(1000.0 * w), so the same w gives the same node coordinates bit for
// bit. Vanadium is absent there, so it weighs 0 there and here, as do
// elements 113 and 115-118.
struct Element {
    const char *symbol;
    double kg_per_mol;
};
constexpr Element kElements[] = {
    {"H", 0.001007940},   {"He", 0.004002602}, {"Li", 0.006941000}, {"Be", 0.009012180},
    {"B", 0.010811000},   {"C", 0.012011000},  {"N", 0.014006740},  {"O", 0.015999400},
    {"F", 0.018998403},   {"Ne", 0.020179700}, {"Na", 0.022989768}, {"Mg", 0.024305000},
    {"Al", 0.026981539},  {"Si", 0.028085500}, {"P", 0.030973762},  {"S", 0.032066000},
    {"Cl", 0.035452700},  {"Ar", 0.039948000}, {"K", 0.039098300},  {"Ca", 0.040078000},
    {"Sc", 0.044955910},  {"Ti", 0.047880000}, {"Cr", 0.051996100},
    {"Mn", 0.054938050},  {"Fe", 0.055847000}, {"Co", 0.058933200}, {"Ni", 0.058693400},
    {"Cu", 0.063546000},  {"Zn", 0.065390000}, {"Ga", 0.069723000}, {"Ge", 0.072610000},
    {"As", 0.074921590},  {"Se", 0.078960000}, {"Br", 0.079904000}, {"Kr", 0.083800000},
    {"Rb", 0.085467800},  {"Sr", 0.087620000}, {"Y", 0.088905850},  {"Zr", 0.091224000},
    {"Nb", 0.092906380},  {"Mo", 0.095940000}, {"Tc", 0.097907200}, {"Ru", 0.101070000},
    {"Rh", 0.102905500},  {"Pd", 0.106420000}, {"Ag", 0.107868200}, {"Cd", 0.112411000},
    {"In", 0.114818000},  {"Sn", 0.118710000}, {"Sb", 0.121760000}, {"Te", 0.127600000},
    {"I", 0.126904470},   {"Xe", 0.131290000}, {"Cs", 0.132905430}, {"Ba", 0.137327000},
    {"La", 0.138905500},  {"Ce", 0.140115000}, {"Pr", 0.140907650}, {"Nd", 0.144240000},
    {"Pm", 0.144912700},  {"Sm", 0.150360000}, {"Eu", 0.151965000}, {"Gd", 0.157250000},
    {"Tb", 0.158925340},  {"Dy", 0.162500000}, {"Ho", 0.164930320}, {"Er", 0.167260000},
    {"Tm", 0.168934210},  {"Yb", 0.173040000}, {"Lu", 0.174967000}, {"Hf", 0.178490000},
    {"Ta", 0.180947900},  {"W", 0.183840000},  {"Re", 0.186207000}, {"Os", 0.190230000},
    {"Ir", 0.192220000},  {"Pt", 0.195080000}, {"Au", 0.196966540}, {"Hg", 0.200590000},
    {"Tl", 0.204383300},  {"Pb", 0.207200000}, {"Bi", 0.208980370}, {"Po", 0.208982400},
    {"At", 0.209987100},  {"Rn", 0.222017600}, {"Fr", 0.223019700}, {"Ra", 0.226025400},
    {"Ac", 0.227027800},  {"Th", 0.232038100}, {"Pa", 0.231035880}, {"U", 0.238028900},
    {"Np", 0.237048000},  {"Pu", 0.244064200}, {"Am", 0.243061400}, {"Cm", 0.247070300},
    {"Bk", 0.247070300},  {"Cf", 0.251079600}, {"Es", 0.252083000}, {"Fm", 0.257095100},
    {"Md", 0.258100000},  {"No", 0.259100900}, {"Lr", 0.262110000}, {"Rf", 0.261000000},
    {"Db", 0.262000000},  {"Sg", 0.266000000}, {"Bh", 0.264000000}, {"Hs", 0.269000000},
    {"Mt", 0.268000000},  {"Ds", 0.269000000}, {"Rg", 0.272000000}, {"Cn", 0.277000000},
    {"Nh", 0.000000000},  {"Fl", 0.289000000}, {"Mc", 0.000000000}, {"Lv", 0.000000000},
    {"Ts", 0.000000000},  {"Og", 0.000000000},
};

// All whitespace removed (utils.cpp removeWhiteSpace).
std::string squeeze(const std::string &s) {
    std::string out;
    out.reserve(s.size());
    for (char c : s)
        if (!std::isspace(static_cast<unsigned char>(c))) out += c;
    return out;
}

int to_int(const std::string &field, const std::string &line) {
    try {
        return std::stoi(field);
    } catch (const std::exception &) {
        throw std::invalid_argument("PDB atom record: bad integer field '" + field + "' in: " +
                                    line);
    }
}

double to_double(const std::string &field, const std::string &line) {
    try {
        return std::stod(field);
    } catch (const std::exception &) {
        throw std::invalid_argument("PDB atom record: bad number field '" + field + "' in: " +
                                    line);
    }
}

uint32_t bswap32(uint32_t v) {
    return (v >> 24) | ((v >> 8) & 0xff00u) | ((v << 8) & 0xff0000u) | (v << 24);
}

// A DCD file: Fortran unformatted records, each framed by its byte length
// (int32) before and after.
class DcdFile {
public:
    explicit DcdFile(const std::string &path) : path_(path), in_(path, std::ios::binary) {
        if (!in_) throw std::runtime_error("cannot open DCD file " + path);
        in_.seekg(0, std::ios::end);
        size_ = (long long)in_.tellg();
        in_.seekg(0);
        uint32_t first = 0;
        read_raw(&first, 4);
        if (first == 84) lay_.swap = false;
        else if (bswap32(first) == 84) lay_.swap = true;
        else throw std::runtime_error(path + ": not a DCD file (first record is not 84 bytes)");
        char magic[4];
        read_raw(magic, 4);
        if (std::memcmp(magic, "CORD", 4) != 0)
            throw std::runtime_error(path + ": not a coordinate DCD file (no CORD tag)");
        int32_t ic[20];
        read_raw(ic, sizeof(ic));
        for (auto &v : ic) v = (int32_t)word((uint32_t)v);
        expect_marker(84);
        const bool charmm = ic[19] != 0;
        lay_.unit_cell = charmm && ic[10] != 0;
        lay_.fourth_dim = charmm && ic[11] != 0;
        if (ic[8] != 0)
            throw std::runtime_error(path + ": DCD files with fixed atoms are not supported");
        const uint32_t title = next_marker();  // title record: skipped
        in_.seekg(title, std::ios::cur);
        expect_marker(title);
        expect_marker(4);
        int32_t natoms = 0;
        read_raw(&natoms, 4);
        natoms = (int32_t)word((uint32_t)natoms);
        expect_marker(4);
        if (natoms <= 0) throw std::runtime_error(path + ": DCD file holds no atoms");
        lay_.natoms = natoms;
        lay_.header_bytes = (long long)in_.tellg();
        const long long coord = 8 + 4LL * natoms;
        lay_.frame_bytes = (lay_.unit_cell ? 8 + 48 : 0) + 3 * coord + (lay_.fourth_dim ? coord : 0);
        lay_.nframes = (int)((size_ - lay_.header_bytes) / lay_.frame_bytes);
    }

    const DcdLayout &layout() const { return lay_; }

    void frame(int f, float *x, float *y, float *z) {
        if (f < 0 || f >= lay_.nframes)
            throw std::out_of_range(path_ + ": frame " + std::to_string(f) + " of " +
                                    std::to_string(lay_.nframes));
        in_.seekg(lay_.header_bytes + (long long)f * lay_.frame_bytes);
        if (lay_.unit_cell) {
            expect_marker(48);
            in_.seekg(48, std::ios::cur);
            expect_marker(48);
        }
        coords(x);
        coords(y);
        coords(z);
    }

private:
    uint32_t word(uint32_t v) const { return lay_.swap ? bswap32(v) : v; }

    void read_raw(void *dst, size_t n) {
        in_.read(static_cast<char *>(dst), (std::streamsize)n);
        if (!in_) throw std::runtime_error(path_ + ": truncated DCD file");
    }

    uint32_t next_marker() {
        uint32_t m = 0;
        read_raw(&m, 4);
        return word(m);
    }

    void expect_marker(uint32_t want) {
        const uint32_t got = next_marker();
        if (got != want)
            throw std::runtime_error(path_ + ": corrupt DCD record (length " +
                                     std::to_string(got) + ", expected " +
                                     std::to_string(want) + ")");
    }

    void coords(float *dst) {
        const uint32_t bytes = 4u * (uint32_t)lay_.natoms;
        expect_marker(bytes);
        read_raw(dst, bytes);
        if (lay_.swap) {
            auto *w = reinterpret_cast<uint32_t *>(dst);
            for (int a = 0; a < lay_.natoms; ++a) w[a] = bswap32(w[a]);
        }
        expect_marker(bytes);
    }

    std::string path_;
    std::ifstream in_;
    long long size_ = 0;
    DcdLayout lay_;
};

}  // namespace

double atomic_weight(const std::string &symbol) {
    static const std::unordered_map<std::string, double> table = [] {
        std::unordered_map<std::string, double> t;
        for (const auto &e : kElements) t.emplace(e.symbol, e.kg_per_mol);
        return t;
    }();
    auto it = table.find(symbol);
    return it == table.end() ? 0.0 : it->second;
}

// atom.cpp:121-199 (the constructor Network::parsePdb uses): fixed columns,
// whitespace removed; serial = index + 1; the temperature factor is read from
// the occupancy columns 55-60, as there.
Atom parse_atom_record(const std::string &line, int index) {
    if (line.size() < 76)
        throw std::invalid_argument("PDB atom record shorter than 76 columns: " + line);
    Atom a;
    a.index = index;
    a.serial = index + 1;
    a.name = squeeze(line.substr(12, 3));
    a.residueName = squeeze(line.substr(17, 3));
    a.chainId = squeeze(line.substr(21, 1));
    const std::string resid = squeeze(line.substr(22, 4));
    a.residueId = to_int(resid, line);
    a.occupancy = (float)to_double(squeeze(line.substr(54, 6)), line);
    a.temperatureFactor = (float)to_double(squeeze(line.substr(54, 6)), line);
    a.segmentId = squeeze(line.substr(72, 4));
    a.element = squeeze(line.substr(76, 2));
    for (auto &c : a.element) c = (char)std::tolower(static_cast<unsigned char>(c));
    if (!a.element.empty()) a.element[0] = (char)std::toupper(static_cast<unsigned char>(a.element[0]));
    a.tag = a.residueName + "_" + resid + "_" + a.chainId + "_" + a.segmentId;
    a.mass = (float)(1000.0 * atomic_weight(a.element));
    return a;
}

void Atoms::add(Atom a) {
    if (std::find(tags_seen_.begin(), tags_seen_.end(), a.tag) == tags_seen_.end()) {
        tags_seen_.push_back(a.tag);
        ++unique_tags_;
    }
    atoms_.push_back(std::make_unique<Atom>(std::move(a)));
}

Atom *Atoms::at(int i) {
    if (i < 0 || i >= numAtoms()) throw std::out_of_range("atom index out of range");
    return atoms_[i].get();
}

std::vector<Atom *> Atoms::atoms() {
    std::vector<Atom *> v;
    v.reserve(atoms_.size());
    for (auto &a : atoms_) v.push_back(a.get());
    return v;
}

DcdLayout read_dcd_layout(const std::string &path) { return DcdFile(path).layout(); }

Network::Network() : coords_(std::make_unique<CuArray<float>>()) {}

// utils.cpp isRecordAtom: any line containing "ATOM" is an atom record.
void Network::parsePdb(const std::string &path) {
    std::ifstream pdb(path);
    if (!pdb) throw std::runtime_error("cannot open PDB file " + path);
    std::string line;
    int index = 0;
    while (std::getline(pdb, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.find("ATOM") != std::string::npos) atoms_.add(parse_atom_record(line, index++));
    }
}

// network.cpp:29-77 and :97-226. One node per distinct atom tag, numbered in
// order of first appearance; node coordinates per kept frame t are
//     sum over the node's atoms, in atom order, of fl(x * mass)
// accumulated in float from 0, then divided by the node's float total mass.
void Network::init(const std::string &trajectoryFile, const std::string &topologyFile,
                   int firstFrame, int lastFrame, int stride) {
    if (stride < 1) throw std::invalid_argument("stride must be >= 1");
    if (firstFrame < 0) throw std::invalid_argument("firstFrame must be >= 0");
    atoms_ = Atoms();
    nodes_.clear();
    node_of_atom_.clear();
    parsePdb(topologyFile);
    std::unordered_map<std::string, Node *> by_tag;
    for (Atom *a : atoms_.atoms()) {
        auto it = by_tag.find(a->tag);
        if (it == by_tag.end()) {
            nodes_.push_back(std::make_unique<Node>());
            Node *n = nodes_.back().get();
            n->index = (int)nodes_.size() - 1;
            n->tag = a->tag;
            it = by_tag.emplace(a->tag, n).first;
        }
        Node *n = it->second;
        n->atoms.push_back(a);
        n->totalMass += a->mass;
        node_of_atom_.push_back(n);
    }

    DcdFile dcd(trajectoryFile);
    const DcdLayout &lay = dcd.layout();
    if (lay.natoms != atoms_.numAtoms())
        throw std::runtime_error("trajectory has " + std::to_string(lay.natoms) +
                                 " atoms, topology " + std::to_string(atoms_.numAtoms()));
    const int last = lastFrame == -1 ? lay.nframes - 1 : lastFrame;
    if (last >= lay.nframes)
        throw std::out_of_range("lastFrame " + std::to_string(last) + " beyond the trajectory's " +
                                std::to_string(lay.nframes) + " frames");
    if (firstFrame > last) throw std::invalid_argument("firstFrame > lastFrame");
    numFrames_ = (last - firstFrame) / stride + 1;

    const int N = numNodes(), F = numFrames_, A = lay.natoms;
    coords_ = std::make_unique<CuArray<float>>();
    coords_->init(N, 3 * F);
    float *c = coords_->host();
    std::vector<float> x(A), y(A), z(A), mass(A);
    std::vector<size_t> row_of(A);
    for (int a = 0; a < A; ++a) {
        mass[a] = atoms_.at(a)->mass;
        row_of[a] = (size_t)node_of_atom_[a]->index * (3 * F);
    }
    for (int t = 0; t < F; ++t) {
        dcd.frame(firstFrame + t * stride, x.data(), y.data(), z.data());
        for (int a = 0; a < A; ++a) {
            float *row = c + row_of[a];
            const float m = mass[a];
            row[t] = row[t] + x[a] * m;
            row[t + F] = row[t + F] + y[a] * m;
            row[t + 2 * F] = row[t + 2 * F] + z[a] * m;
        }
    }
    for (auto &n : nodes_) {
        float *row = c + (size_t)n->index * (3 * F);
        for (int j = 0; j < 3 * F; ++j) row[j] = row[j] / n->totalMass;
    }
}

void Network::nodeCoordinates(const std::string &npyFile) {
    auto fresh = std::make_unique<CuArray<float>>();
    if (fresh->load(npyFile) != 0) throw std::runtime_error("cannot load " + npyFile);
    coords_ = std::move(fresh);
}

std::vector<Node *> Network::nodes() {
    std::vector<Node *> v;
    v.reserve(nodes_.size());
    for (auto &n : nodes_) v.push_back(n.get());
    return v;
}

Node *Network::nodeFromAtomIndex(int atomIndex) {
    if (atomIndex < 0 || atomIndex >= (int)node_of_atom_.size())
        throw std::out_of_range("atom index out of range");
    return node_of_atom_[atomIndex];
}

}  // namespace netchem
