// CircuitFile.h - binary circuit files (.qcsc) for C++ callers: the format
// qcsim_amd/circuit.py writes and reads, little-endian:
//   "QCSC", uint32 version (1), uint32 qubits, uint32 gate count,
//   then per gate: uint32 k, traits, flags, qubits[5], 2 * 4^k float64
//   (the 2^k x 2^k complex matrix, row-major, re then im).
// Records map one to one onto qcs_gate (include/qcsim_amd.h), so a file can
// go straight to qcs_apply_circuit, qcs_fuse or qcs_program_create. A
// reference AppliedGate list (SimpleGates.h:439-578) is written with
// traits 0 and qubits (q1, q2, q3).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../qcsim_amd.h"

namespace QC {

struct CircuitFile {
    size_t nrQubits = 0;
    std::vector<qcs_gate> gates;
};

inline CircuitFile ReadCircuitFile(const std::string& path) {
    std::FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("cannot open " + path);
    CircuitFile c;
    auto fail = [&](const char* what) {
        std::fclose(f);
        throw std::runtime_error(std::string(what) + ": " + path);
    };
    char magic[4];
    uint32_t hdr[3];
    if (std::fread(magic, 1, 4, f) != 4 || std::memcmp(magic, "QCSC", 4) != 0) fail("not a circuit file");
    if (std::fread(hdr, sizeof(uint32_t), 3, f) != 3) fail("truncated header");
    if (hdr[0] != 1) fail("unsupported circuit file version");
    c.nrQubits = hdr[1];
    c.gates.resize(hdr[2]);
    for (auto& g : c.gates) {
        uint32_t rec[8];
        if (std::fread(rec, sizeof(uint32_t), 8, f) != 8) fail("truncated gate record");
        std::memset(&g, 0, sizeof g);
        g.k = rec[0];
        g.traits = rec[1];
        g.flags = rec[2];
        if (g.k < 1 || g.k > 5) fail("bad gate arity");
        for (int i = 0; i < 5; ++i) g.qubits[i] = rec[3 + i];
        for (uint32_t i = 0; i < g.k; ++i)
            if (g.qubits[i] >= c.nrQubits) fail("qubit out of range");
        const size_t nd = size_t(2) << (2 * g.k);
        if (std::fread(g.mat, sizeof(double), nd, f) != nd) fail("truncated matrix");
    }
    if (std::fgetc(f) != EOF) fail("trailing bytes");
    std::fclose(f);
    return c;
}

inline void WriteCircuitFile(const std::string& path, const CircuitFile& c) {
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot create " + path);
    const uint32_t hdr[3] = {1u, uint32_t(c.nrQubits), uint32_t(c.gates.size())};
    bool ok = std::fwrite("QCSC", 1, 4, f) == 4 && std::fwrite(hdr, sizeof(uint32_t), 3, f) == 3;
    for (const auto& g : c.gates) {
        if (!ok) break;
        const uint32_t rec[8] = {g.k, g.traits, g.flags, g.qubits[0], g.qubits[1], g.qubits[2], g.qubits[3],
                                 g.qubits[4]};
        const size_t nd = size_t(2) << (2 * g.k);
        ok = std::fwrite(rec, sizeof(uint32_t), 8, f) == 8 && std::fwrite(g.mat, sizeof(double), nd, f) == nd;
    }
    if (std::fclose(f) != 0 || !ok) throw std::runtime_error("write failed: " + path);
}

}  // namespace QC
