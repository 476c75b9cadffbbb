// Particle file I/O: the FastASCII / ASCII formats of the reference
// (aquagpusph/InputOutput/ASCII.cpp formatLine + FastASCII.cpp readField):
// '#' starts a comment, blanks and ";()[]{}" act as commas, each field
// consumes as many numbers as its type has components.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "sim.hpp"

namespace aq {
namespace rt {

void load_fastascii(Sim& S, const std::string& path, size_t offset, size_t n,
                    const std::vector<std::string>& fields)
{
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) throw Error(AQ_ENOTFOUND, "cannot open particles file '" + path + "'");
    std::vector<Var*> vs;
    size_t ncols = 0;
    for (auto& name : fields) {
        Var& v = S.vars.at(name);
        if (!v.is_array())
            throw Error(AQ_EINVAL, "cannot load the scalar '" + name + "'");
        if (v.length < offset + n)
            throw Error(AQ_EINVAL, "variable '" + name + "' is too short");
        if (v.type.base == 'd')
            throw Error(AQ_EUNSUP, "cannot load the double field '" + name +
                                       "' (FastASCII loads 32-bit fields)");
        vs.push_back(&v);
        ncols += v.type.ncomp;
    }
    std::vector<std::vector<uint32_t>> data(vs.size());
    for (size_t k = 0; k < vs.size(); k++)
        data[k].assign(n * vs[k]->type.stored, 0u);
    size_t row = 0, lineno = 0;
    std::string buf;
    while (true) {
        buf.clear();
        int c;
        while ((c = fgetc(f)) != EOF && c != '\n') buf += (char)c;
        if (c == EOF && buf.empty()) break;
        lineno++;
        const size_t hash = buf.find('#');
        if (hash != std::string::npos) buf.erase(hash);
        for (auto& ch : buf)
            if (strchr(" ;()[]{}\t\r", ch)) ch = ',';
        // tokens
        std::vector<const char*> tok;
        char* p = &buf[0];
        const size_t L = buf.size();
        for (size_t i = 0; i < L; i++) {
            if (buf[i] == ',') {
                buf[i] = '\0';
                continue;
            }
            if (i == 0 || buf[i - 1] == '\0') tok.push_back(p + i);
        }
        if (tok.empty()) continue;
        if (tok.size() != ncols) {
            fclose(f);
            throw Error(AQ_EPARSE, path + ":" + std::to_string(lineno) + ": " +
                                       std::to_string(tok.size()) +
                                       " fields, " + std::to_string(ncols) +
                                       " expected");
        }
        if (row >= n) {
            fclose(f);
            throw Error(AQ_EPARSE, path + ": more than " + std::to_string(n) +
                                       " particles");
        }
        size_t t = 0;
        for (size_t k = 0; k < vs.size(); k++) {
            const Type& ty = vs[k]->type;
            for (int cpt = 0; cpt < ty.ncomp; cpt++, t++) {
                uint32_t w;
                if (ty.base == 'f') {
                    const float x = strtof(tok[t], nullptr);
                    memcpy(&w, &x, 4);
                } else if (ty.base == 'i') {
                    const int32_t x = (int32_t)strtol(tok[t], nullptr, 10);
                    memcpy(&w, &x, 4);
                } else {
                    w = (uint32_t)strtoul(tok[t], nullptr, 10);
                }
                data[k][row * ty.stored + cpt] = w;
            }
        }
        row++;
        if (c == EOF) break;
    }
    fclose(f);
    if (row != n)
        throw Error(AQ_EPARSE, path + ": " + std::to_string(row) +
                                   " particles read, " + std::to_string(n) +
                                   " expected");
    for (size_t k = 0; k < vs.size(); k++) {
        const size_t eb = vs[k]->type.elem_bytes();
        AQ_HIP(hipMemcpy((char*)vs[k]->dev + offset * eb, data[k].data(),
                         n * eb, hipMemcpyHostToDevice));
    }
}

}  // namespace rt
}  // namespace aq
