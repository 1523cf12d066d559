// Lipid compositions of the membrane (src/lipid_composition.C:18-333) and
// their packing onto triangles at setg0 (src/interpLib.C:3127-3196,
// 3286-3358).  Host code.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace umcp {

constexpr int kMaxLipidTypes = 64;

// The composition's lipid types (in order of first use), their spontaneous
// curvature (1/A), area per lipid (A^2) and input parts per leaflet, and the
// seed of the rand() stream the packing draws from.
struct LipidComposition {
    std::vector<std::string> names;
    std::vector<double> c0, apl, inner, outer;
    unsigned int seed = 1;
    int ntypes() const { return (int)c0.size(); }
};

// glibc's rand() (random_r's TYPE_3 additive feedback generator): srand(seed)
// then rand(), in [0, 2^31 - 1]
class GlibcRand {
public:
    explicit GlibcRand(unsigned int seed);
    int next();

private:
    int32_t r_[31];
    int f_ = 3, b_ = 0;
};

// The reference's lipid library extended by a composition file's "lipid
// library NAME APL c0 [atom]" lines, and the composition of its "lipid inner
// NAME parts" / "lipid outer NAME parts" lines (case-insensitive, types in
// order of first mention; a leaflet without parts gets one part of the first
// type; no inner/outer lines: one part of "ZERO" in each leaflet).  0 or -1
// with err set.
int read_lipid_composition(const char *path, unsigned int seed, LipidComposition &out, std::string &err);

// Packs one triangle of reference area A: leaflet fractions from the input
// parts, draws of rand()/RAND_MAX pick a type by the cumulative fractions and
// add its area per lipid until a leaflet holds more than A - 65/2.  Regular
// faces draw one (inner, outer) pair per round, irregular faces one pair per
// type within a round, as the reference does.  inner/outer receive each
// type's area (ntypes each).
void pack_lipids(const LipidComposition &L, GlibcRand &g, double A, bool irregular, double *inner, double *outer);

}  // namespace umcp
