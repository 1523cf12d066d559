// Lipid compositions and their packing (see umcp_lipids.hpp).
//
// Reference: alexsodt/UMCP src/lipid_composition.C:18-34 (library), 47-333
// (readLipidComposition), src/interpLib.C:3127-3196 and 3286-3358 (setg0's
// rand()-driven packing of regular and irregular faces).
#include "umcp_lipids.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <strings.h>

namespace umcp {

GlibcRand::GlibcRand(unsigned int seed)
{
    if (seed == 0) seed = 1;
    int32_t word = (int32_t)seed;   // glibc keeps the seed in 32 signed bits
    r_[0] = word;
    for (int i = 1; i < 31; i++) {
        // 16807 word mod (2^31 - 1) without overflow (Schrage)
        const int64_t hi = word / 127773, lo = word % 127773;
        word = (int32_t)(16807 * lo - 2836 * hi);
        if (word < 0) word += 2147483647;
        r_[i] = word;
    }
    for (int i = 0; i < 310; i++) (void)next();
}

int GlibcRand::next()
{
    const uint32_t v = (uint32_t)r_[f_] + (uint32_t)r_[b_];
    r_[f_] = (int32_t)v;
    f_ = (f_ + 1) % 31;
    b_ = (b_ + 1) % 31;
    return (int)(v >> 1);
}

namespace {

struct LibLipid {
    std::string name;
    double c0, apl;
};

// the fixed library (Chen/Rand 1997; DOPS: Fuller 2003, area from the
// CHARMM force field)
std::vector<LibLipid> fixed_library()
{
    return {{"ZERO", 0.0, 68.3},          {"POPC", -1.0 / 87.3, 68.3}, {"DOPE", -1.0 / 29.4, 63.4},
            {"POPE", -1.0 / 29.4, 58.8},  {"DOPC", -1.0 / 87.3, 69.7}, {"CHOL", -1.0 / 27.2, 31.0},
            {"CHL1", -1.0 / 27.2, 31.0},  {"DOG", -1.0 / 11.5, 59.0},  {"DOPS", 1.0 / 144.0, 71.6},
            {"CER180", 0.0, 59.0},        {"SSM", 0.0, 59.0},          {"PSM", 0.0, 59.0}};
}

int find_name(const std::vector<LibLipid> &lib, const char *name)
{
    int found = -1;
    for (size_t i = 0; i < lib.size(); i++)
        if (!strcasecmp(lib[i].name.c_str(), name)) found = (int)i;   // the last match, as the reference
    return found;
}

}  // namespace

int read_lipid_composition(const char *path, unsigned int seed, LipidComposition &out, std::string &err)
{
    FILE *fp = fopen(path, "r");
    if (!fp) {
        err = std::string("cannot open ") + path;
        return -1;
    }
    std::vector<LibLipid> lib = fixed_library();
    std::vector<std::string> lines;
    char buf[4096];
    while (fgets(buf, sizeof buf, fp)) lines.emplace_back(buf);
    fclose(fp);
    auto is_lipid = [](const char *s) {
        return !strncasecmp(s, "lipid ", 6) || !strncasecmp(s, "lipid\t", 6);
    };
    // pass 0: library directives
    for (const std::string &l : lines) {
        if (!is_lipid(l.c_str())) continue;
        char junk[256], cmd[256], name[256], atom[256];
        double apl = 0, c0 = 0;
        if (sscanf(l.c_str(), "%255s %255s", junk, cmd) != 2 || strcasecmp(cmd, "library")) continue;
        const int nr = sscanf(l.c_str(), "%255s %255s %255s %lf %lf %255s", junk, cmd, name, &apl, &c0, atom);
        if (nr < 5) {
            err = "expected 'lipid library lipidName APL c0': " + l;
            return -1;
        }
        if (!(apl > 0.0) || !std::isfinite(apl) || !std::isfinite(c0)) {
            err = "a library lipid needs a finite area per lipid above 0 and a finite c0: " + l;
            return -1;
        }
        const int at = find_name(lib, name);
        if (at < 0)
            lib.push_back({name, c0, apl});
        else
            lib[at].c0 = c0, lib[at].apl = apl;
    }
    // pass 1: the composition's types in order of first mention; pass 2: parts
    LipidComposition L;
    L.seed = seed;
    for (int pass = 1; pass <= 2; pass++) {
        for (const std::string &l : lines) {
            if (!is_lipid(l.c_str())) continue;
            char junk[256], cmd[256], name[256];
            double parts = 0;
            if (sscanf(l.c_str(), "%255s %255s", junk, cmd) != 2) continue;
            const bool inner = !strcasecmp(cmd, "inner"), outer = !strcasecmp(cmd, "outer");
            if (!inner && !outer) continue;
            if (sscanf(l.c_str(), "%255s %255s %255s %lf", junk, cmd, name, &parts) != 4) {
                err = std::string("expected 'lipid ") + cmd + " lipidName parts': " + l;
                return -1;
            }
            const int at = find_name(lib, name);
            if (at < 0) {
                err = std::string("lipid '") + name + "' is not in the library";
                return -1;
            }
            int x = -1;
            for (int i = 0; i < L.ntypes(); i++)
                if (!strcasecmp(L.names[i].c_str(), name)) x = i;
            if (pass == 1) {
                if (x < 0) {
                    if (L.ntypes() == kMaxLipidTypes) {
                        err = "more than 64 lipid types";
                        return -1;
                    }
                    L.names.push_back(lib[at].name);
                    L.c0.push_back(lib[at].c0);
                    L.apl.push_back(lib[at].apl);
                    L.inner.push_back(0.0);
                    L.outer.push_back(0.0);
                }
            } else {
                (inner ? L.inner : L.outer)[x] += parts;
            }
        }
    }
    if (L.ntypes() == 0) {
        L.names.push_back("ZERO");
        L.c0.push_back(0.0);
        L.apl.push_back(68.3);
        L.inner.push_back(0.0);
        L.outer.push_back(0.0);
    }
    double si = 0, so = 0;
    for (int x = 0; x < L.ntypes(); x++) si += L.inner[x], so += L.outer[x];
    if (so == 0) L.outer[0] = 1;
    if (si == 0) L.inner[0] = 1;
    out = L;
    return 0;
}

void pack_lipids(const LipidComposition &L, GlibcRand &g, double A, bool irregular, double *inner, double *outer)
{
    const int n = L.ntypes();
    double fr_i[kMaxLipidTypes], fr_o[kMaxLipidTypes];
    double sum_i = 0.0, sum_o = 0.0;
    for (int x = 0; x < n; x++) {
        inner[x] = outer[x] = 0.0;
        fr_i[x] = L.inner[x];
        fr_o[x] = L.outer[x];
        sum_i += fr_i[x];
        sum_o += fr_o[x];
    }
    for (int x = 0; x < n; x++) {
        fr_i[x] /= sum_i;
        fr_o[x] /= sum_o;
    }
    for (int x = 1; x < n; x++) fr_i[x] += fr_i[x - 1];
    for (int x = 1; x < n; x++) fr_o[x] += fr_o[x - 1];
    const double area_per_lipid_av = 65.0;
    const double rand_max = 2147483647.0;
    double a_i = 0.0, a_o = 0.0;
    for (;;) {
        double rn_i = 0.0, rn_o = 0.0;
        if (!irregular) {
            rn_i = g.next() / rand_max;
            rn_o = g.next() / rand_max;
        }
        const bool idone0 = a_i > 0 && a_i > A - 0.5 * area_per_lipid_av;
        const bool odone0 = a_o > 0 && a_o > A - 0.5 * area_per_lipid_av;
        if (idone0 && odone0) break;
        bool idone = idone0, odone = odone0;
        for (int x = 0; x < n; x++) {
            if (irregular) {
                rn_i = g.next() / rand_max;
                rn_o = g.next() / rand_max;
            }
            if (rn_i < fr_i[x] && !idone) {
                inner[x] += L.apl[x];
                idone = true;
                a_i += L.apl[x];
            }
            if (rn_o < fr_o[x] && !odone) {
                outer[x] += L.apl[x];
                odone = true;
                a_o += L.apl[x];
            }
        }
    }
}

}  // namespace umcp
