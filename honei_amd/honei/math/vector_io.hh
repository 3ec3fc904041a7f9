// honei/math/vector_io.hh -- VectorIO<EXP> text and VectorIO<DV> binary
// formats (math/vector_io.hh:52-250).
#pragma once

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "honei/la/dense_vector.hh"

namespace honei {

template <typename IOFormat_> struct VectorIO;

template <> struct VectorIO<io_formats::EXP> {
    // one value per line; lines containing '#' are skipped
    template <typename DT>
    static DenseVector<DT> read_vector(const std::string &path, DT) {
        std::ifstream in(path.c_str());
        if (!in.is_open()) throw InternalError("Unable to open Vector file " + path);
        std::vector<DT> data;
        std::string line;
        while (std::getline(in, line)) {
            if (line.find('#') != std::string::npos) continue;
            const size_t p = line.find_first_not_of(" \t\r");
            if (p == std::string::npos) continue;
            data.push_back((DT)std::atof(line.c_str() + p));
        }
        DenseVector<DT> r(data.size());
        for (size_t i = 0; i < data.size(); ++i) r.elements()[i] = data[i];
        return r;
    }
};

template <> struct VectorIO<io_formats::DV> {
    template <typename DT>
    static DenseVector<DT> read_vector(const std::string &path, DT) {
        FILE *f = std::fopen(path.c_str(), "rb");
        if (!f) throw InternalError("File " + path + " not found!");
        uint64_t n = 0;
        if (std::fread(&n, sizeof(uint64_t), 1, f) != 1) {
            std::fclose(f);
            throw InternalError("fread error!");
        }
        std::vector<double> v(n);
        const bool ok = std::fread(v.data(), sizeof(double), n, f) == n;
        std::fclose(f);
        if (!ok) throw InternalError("fread error!");
        DenseVector<DT> r(n);
        for (uint64_t i = 0; i < n; ++i) r.elements()[i] = (DT)v[i];
        return r;
    }
};

}  // namespace honei
