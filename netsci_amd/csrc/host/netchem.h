// netchem: the producer of the pair loop's input. Reads a PDB topology and a
// CHARMM/NAMD DCD trajectory and forms one node per residue tag with its
// mass-weighted centre of coordinates per frame, laid out as the KSG path
// expects (row = node, columns = [x frames | y frames | z frames]).
//
// Reference: netchem/src/network.cpp:29-255 (Network::init, parseDcd,
// parsePdb), netchem/src/atom.cpp:51-153 (PDB columns, tag, mass),
// netchem/include/chemical_properties.h (atomic weights),
// netchem/src/utils.cpp (record test, whitespace removal). The DCD reader is
// this project's own, written from the published CHARMM/NAMD format; the
// reference vendors VMD's dcdplugin (netchem/src/dcd/dcd.cpp).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "cuarray.h"

namespace netchem {

// Standard atomic weight in kg/mol of an element symbol ("C", "Fe"); 0 for
// a symbol the table does not hold.
double atomic_weight(const std::string &symbol);

struct Atom {
    int index = 0;   // position among the topology's atom records
    int serial = 0;  // index + 1 (the record's own serial is not used)
    std::string name, element, residueName, chainId, segmentId, tag;
    int residueId = 0;
    float occupancy = 0.0f;
    float temperatureFactor = 0.0f;
    float mass = 0.0f;  // (float)(1000 * kg/mol), i.e. g/mol
};

// One atom record of a PDB file (the reference's column rules).
Atom parse_atom_record(const std::string &line, int index);

class Atoms {
public:
    Atoms() = default;
    Atoms(const Atoms &) = delete;
    Atoms &operator=(const Atoms &) = delete;
    Atoms(Atoms &&) = default;
    Atoms &operator=(Atoms &&) = default;
    void add(Atom a);
    int numAtoms() const { return (int)atoms_.size(); }
    int numUniqueTags() const { return unique_tags_; }
    Atom *at(int i);
    std::vector<Atom *> atoms();

private:
    std::vector<std::unique_ptr<Atom>> atoms_;
    std::vector<std::string> tags_seen_;
    int unique_tags_ = 0;
};

struct Node {
    int index = 0;
    std::string tag;
    std::vector<Atom *> atoms;
    float totalMass = 0.0f;  // float sum of the atoms' masses in atom order
    int numAtoms() const { return (int)atoms.size(); }
};

// Frame layout of a DCD file, from its header.
struct DcdLayout {
    bool swap = false;        // file written with the other byte order
    bool unit_cell = false;   // a 6-double unit-cell record precedes each frame
    bool fourth_dim = false;  // a fourth coordinate record follows each frame
    int natoms = 0;
    int nframes = 0;          // whole frames the file holds
    long long header_bytes = 0;
    long long frame_bytes = 0;
};

DcdLayout read_dcd_layout(const std::string &path);

class Network {
public:
    Network();
    Network(const Network &) = delete;
    Network &operator=(const Network &) = delete;
    // Frames firstFrame, firstFrame + stride, ... <= lastFrame (lastFrame =
    // -1: the trajectory's last frame).
    void init(const std::string &trajectoryFile, const std::string &topologyFile,
              int firstFrame, int lastFrame, int stride = 1);
    int numNodes() const { return (int)nodes_.size(); }
    int numFrames() const { return numFrames_; }
    CuArray<float> *nodeCoordinates() { return coords_.get(); }
    // Replace the coordinates with the contents of a .npy file.
    void nodeCoordinates(const std::string &npyFile);
    std::vector<Node *> nodes();
    Atoms *atoms() { return &atoms_; }
    Node *nodeFromAtomIndex(int atomIndex);

private:
    void parsePdb(const std::string &path);

    Atoms atoms_;
    std::vector<std::unique_ptr<Node>> nodes_;
    std::vector<Node *> node_of_atom_;
    std::unique_ptr<CuArray<float>> coords_;
    int numFrames_ = 0;
};

}  // namespace netchem
