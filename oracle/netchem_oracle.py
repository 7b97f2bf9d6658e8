"""CPU oracle for the netchem producer — test infrastructure only.

A numpy restatement, independent of netsci_amd/csrc/host/netchem.cpp, of how
the reference turns a PDB topology and a DCD trajectory into node
coordinates. Used as the checker by tests/; product code never imports it.

* atom records: netchem/src/utils.cpp:47-55 (a line containing "ATOM"),
  netchem/src/atom.cpp:121-199 (columns, tag, element case, mass);
* nodes and coordinates: netchem/src/network.cpp:29-77 (nodes numbered by
  first appearance of a tag, total mass summed in float in atom order) and
  :97-226 (per kept frame, fl(acc + fl(x * mass)) over the atoms in index
  order from 0, then one float division by the node's total mass);
* DCD: the CHARMM/NAMD layout (Fortran records framed by their byte length;
  an 84-byte CORD header whose control word 19 marks CHARMM files and word
  10 a per-frame unit-cell record; a title record; an atom-count record; per
  frame X, Y, Z records of float32), parsed here with struct/numpy.

Pinned by: the reference's netchem tests (290 nodes, 10 frames, stride
equalities), DCD frame 0 of tests/golden/netchem/test.dcd against the
coordinates printed in test.pdb (3 decimals), standard atomic weights, and a
hand-computed two-atom node.
"""
import struct

import numpy as np

# kg/mol of the elements that occur in biomolecular topologies, as in
# netchem/include/chemical_properties.h (standard atomic weights).
KG_PER_MOL = {"H": 0.001007940, "C": 0.012011000, "N": 0.014006740, "O": 0.015999400,
              "S": 0.032066000, "P": 0.030973762, "Na": 0.022989768, "Cl": 0.035452700,
              "Mg": 0.024305000, "K": 0.039098300, "Ca": 0.040078000, "Fe": 0.055847000,
              "Zn": 0.065390000}


def atom_records(pdb_path):
    """(tag, mass float32) per atom record, in file order."""
    out = []
    for line in open(pdb_path):
        line = line.rstrip("\n").rstrip("\r")
        if "ATOM" not in line:
            continue
        squeeze = lambda s: "".join(s.split())  # noqa: E731
        element = squeeze(line[76:78]).lower()
        element = element[:1].upper() + element[1:]
        tag = "_".join([squeeze(line[17:20]), squeeze(line[22:26]), squeeze(line[21:22]),
                        squeeze(line[72:76])])
        mass = np.float32(1000.0 * KG_PER_MOL.get(element, 0.0))
        out.append((tag, mass))
    return out


def dcd_frames(dcd_path):
    """All frames of a little- or big-endian DCD file: float32 (F, 3, natoms)."""
    b = open(dcd_path, "rb").read()
    end = "<" if struct.unpack("<i", b[:4])[0] == 84 else ">"
    assert struct.unpack(end + "i", b[:4])[0] == 84 and b[4:8] == b"CORD"
    ic = struct.unpack(end + "20i", b[8:88])
    assert ic[8] == 0, "fixed atoms"
    cell = ic[19] != 0 and ic[10] != 0
    fourth = ic[19] != 0 and ic[11] != 0
    off = 92
    (lt,) = struct.unpack(end + "i", b[off:off + 4])
    off += 4 + lt + 4
    (_, natom, _) = struct.unpack(end + "3i", b[off:off + 12])
    off += 12
    frame_bytes = (56 if cell else 0) + (3 + int(fourth)) * (8 + 4 * natom)
    frames = []
    while off + frame_bytes <= len(b):
        if cell:
            off += 56
        xyz = []
        for _ in range(3):
            (n,) = struct.unpack(end + "i", b[off:off + 4])
            assert n == 4 * natom
            xyz.append(np.frombuffer(b, end + "f4", natom, off + 4).astype(np.float32))
            off += 8 + n
        if fourth:
            off += 8 + 4 * natom
        frames.append(np.stack(xyz))
    return np.array(frames, dtype=np.float32)


def node_coordinates(dcd_path, pdb_path, first, last, stride=1):
    """(N, 3F) float32 node coordinates and the node tags, in the reference's
    float arithmetic and order."""
    atoms = atom_records(pdb_path)
    tags, node_of = [], []
    for tag, _ in atoms:
        if tag not in tags:
            tags.append(tag)
        node_of.append(tags.index(tag))
    N = len(tags)
    total = np.zeros(N, np.float32)
    for (_, m), n in zip(atoms, node_of):
        total[n] = np.float32(total[n] + m)
    frames = dcd_frames(dcd_path)
    if last == -1:
        last = len(frames) - 1
    keep = list(range(first, last + 1, stride))
    F = len(keep)
    acc = np.zeros((N, 3, F), np.float32)
    for t, f in enumerate(keep):
        xyz = frames[f]
        for a, ((_, m), n) in enumerate(zip(atoms, node_of)):
            acc[n, :, t] = (acc[n, :, t] + (xyz[:, a] * m).astype(np.float32)).astype(np.float32)
    out = (acc / total[:, None, None]).astype(np.float32)
    return out.reshape(N, 3 * F), tags
