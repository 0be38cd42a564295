// AMBER prmtop / inpcrd readers and hydrogen-bond promotion.
//
// prmtop::parse follows geodesic-io/src/prmtop.rs:14-344 (paths relative to
// the reference repository root): %FLAG sections of fixed-width fields,
// sigma = (A/B)^(1/6), epsilon = B^2/(4A) (0 when A or B is 0), 10-12
// H-bond types rejected, bond/angle/dihedral indices as 3x the 0-based atom
// (dihedrals through abs(), dropping the 1-4 / improper flags), exclusions
// from NUMBER_EXCLUDED_ATOMS / EXCLUDED_ATOMS_LIST skipping 0.  With
// PrmtopOptions::fidelity the flags are kept instead: every proper dihedral
// whose third index is not negated (1-4 not suppressed) and whose fourth is
// not negated (not an improper) adds its end atoms as a 1-4 pair, scaled by
// 1/SCNB of its dihedral type (SCNB_SCALE_FACTOR; AMBER's default 2.0 when the
// section is absent).  The engine applies one scale to every 1-4 pair, so
// differing SCNB values among the pairs are rejected.  SCEE (electrostatic
// 1-4 scale) is read and kept, but this force field has no electrostatics.
// inpcrd::parse follows inpcrd.rs:12-97; promote_hydrogen_bonds follows
// geodesic-engine/src/constraint.rs:10-40.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <map>
#include <set>
#include <sstream>

#include "geodesic.hpp"

namespace geodesic {
namespace {

constexpr double kAmberChargeToElementary = 18.2223;  // prmtop.rs:9
constexpr double kAmberVelocityToAngPerPs = 20.455;   // inpcrd.rs:7

// str::lines(): split on '\n', dropping one trailing '\r' per line.
std::vector<std::string> lines_of(const std::string &text) {
    std::vector<std::string> out;
    size_t b = 0;
    while (b < text.size()) {
        size_t e = text.find('\n', b);
        if (e == std::string::npos) e = text.size();
        std::string l = text.substr(b, e - b);
        if (!l.empty() && l.back() == '\r') l.pop_back();
        out.push_back(l);
        b = e + 1;
    }
    return out;
}

std::string trim(const std::string &s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) b++;
    while (e > b && std::isspace((unsigned char)s[e - 1])) e--;
    return s.substr(b, e - b);
}

// fixed_width.rs: `count` fields of `width` characters; past the end, "".
std::vector<std::string> chunk_fields(const std::string &raw, size_t width, size_t count) {
    std::vector<std::string> f;
    f.reserve(count);
    for (size_t i = 0; i < count; i++) {
        const size_t start = i * width;
        f.push_back(start < raw.size() ? raw.substr(start, std::min(width, raw.size() - start)) : std::string());
    }
    return f;
}

bool parse_i64(const std::string &field, int64_t &v) {
    const std::string t = trim(field);
    if (t.empty()) return false;
    errno = 0;
    char *end = nullptr;
    const long long r = std::strtoll(t.c_str(), &end, 10);
    if (errno != 0 || end != t.c_str() + t.size()) return false;
    v = r;
    return true;
}

bool parse_f64(const std::string &field, double &v) {
    const std::string t = trim(field);
    if (t.empty()) return false;
    errno = 0;
    char *end = nullptr;
    const double r = std::strtod(t.c_str(), &end);
    if (end != t.c_str() + t.size()) return false;
    v = r;
    return true;
}

// prmtop.rs:14-49
class Sections {
public:
    explicit Sections(const std::string &text) {
        std::string current;
        bool have = false;
        for (const std::string &line : lines_of(text)) {
            if (line.rfind("%FLAG ", 0) == 0) {
                current = trim(line.substr(6));
                have = true;
                map_[current];
                continue;
            }
            if (line.rfind("%FORMAT", 0) == 0) continue;
            if (!line.empty() && line[0] == '%') continue;
            if (have) map_[current] += line;
        }
    }
    bool has(const std::string &flag) const { return map_.count(flag) != 0; }
    const std::string &raw(const std::string &flag) const {
        auto it = map_.find(flag);
        if (it == map_.end()) throw ConfigError("missing required config key: %FLAG " + flag);
        return it->second;
    }
    std::vector<int64_t> ints(const std::string &flag, size_t width, size_t count) const {
        std::vector<int64_t> out;
        out.reserve(count);
        for (const std::string &f : chunk_fields(raw(flag), width, count)) {
            int64_t v = 0;
            if (!parse_i64(f, v))
                throw ConfigError("invalid value for key '%FLAG " + flag + "': '" + f + "' (expected an integer)");
            out.push_back(v);
        }
        return out;
    }
    std::vector<double> floats(const std::string &flag, size_t width, size_t count) const {
        std::vector<double> out;
        out.reserve(count);
        for (const std::string &f : chunk_fields(raw(flag), width, count)) {
            double v = 0;
            if (!parse_f64(f, v))
                throw ConfigError("invalid value for key '%FLAG " + flag + "': '" + f + "' (expected a float)");
            out.push_back(v);
        }
        return out;
    }
    std::vector<std::array<char, 4>> names4(const std::string &flag, size_t count) const {
        std::vector<std::array<char, 4>> out;
        for (const std::string &f : chunk_fields(raw(flag), 4, count)) {
            std::array<char, 4> n{{' ', ' ', ' ', ' '}};
            for (size_t k = 0; k < f.size() && k < 4; k++) n[k] = f[k];
            out.push_back(n);
        }
        return out;
    }

private:
    std::map<std::string, std::string> map_;
};

// prmtop.rs:101-117: the first alphabetic character decides.
Element infer_element(const std::array<char, 4> &name) {
    for (char c : name) {
        if (std::isalpha((unsigned char)c)) {
            switch (std::toupper((unsigned char)c)) {
                case 'H': return Element::H;
                case 'C': return Element::C;
                case 'N': return Element::N;
                case 'O': return Element::O;
                case 'S': return Element::S;
                default: return Element::Unknown;
            }
        }
    }
    return Element::Unknown;
}

// prmtop.rs:119-130
uint32_t residue_id_for_atom(size_t atom0, const std::vector<int64_t> &residue_pointer) {
    const int64_t atom1 = (int64_t)atom0 + 1;
    size_t res = 0;
    for (size_t r = 0; r < residue_pointer.size(); r++) {
        if (residue_pointer[r] <= atom1) res = r;
        else break;
    }
    return (uint32_t)res;
}

// Rust `x as u32` for an f64: NaN and negatives saturate to 0, large to u32::MAX.
uint32_t f64_to_u32(double v) {
    if (!(v > 0.0)) return 0u;
    if (v >= 4294967295.0) return 4294967295u;
    return (uint32_t)v;
}

}  // namespace

namespace prmtop {

std::pair<AtomData, BondedTopology> parse(const std::string &text, const PrmtopOptions &opt) {
    const Sections sec(text);
    const std::vector<int64_t> p = sec.ints("POINTERS", 8, 20);
    const size_t natom = (size_t)p[0], ntypes = (size_t)p[1], nbonh = (size_t)p[2], ntheth = (size_t)p[4];
    const size_t nphih = (size_t)p[6], nnb = (size_t)p[10], nres = (size_t)p[11], nbona = (size_t)p[12];
    const size_t ntheta = (size_t)p[13], nphia = (size_t)p[14], numbnd = (size_t)p[15], numang = (size_t)p[16];
    const size_t nptra = (size_t)p[17];

    const auto names = sec.names4("ATOM_NAME", natom);
    const std::vector<double> charge_raw = sec.floats("CHARGE", 16, natom);
    const std::vector<double> mass = sec.floats("MASS", 16, natom);
    const std::vector<int64_t> type_index = sec.ints("ATOM_TYPE_INDEX", 8, natom);
    const std::vector<int64_t> n_excluded = sec.ints("NUMBER_EXCLUDED_ATOMS", 8, natom);
    const std::vector<int64_t> nb_parm = sec.ints("NONBONDED_PARM_INDEX", 8, ntypes * ntypes);
    const size_t n_lj = ntypes * (ntypes + 1) / 2;
    const std::vector<double> acoef = sec.floats("LENNARD_JONES_ACOEF", 16, n_lj);
    const std::vector<double> bcoef = sec.floats("LENNARD_JONES_BCOEF", 16, n_lj);
    const std::vector<int64_t> residue_pointer = sec.ints("RESIDUE_POINTER", 8, nres);
    const std::vector<int64_t> excluded = sec.ints("EXCLUDED_ATOMS_LIST", 8, nnb);

    AtomData atoms;
    atoms.mass = mass;
    for (double q : charge_raw) atoms.charge.push_back(q / kAmberChargeToElementary);
    for (size_t i = 0; i < natom; i++) {
        const int64_t ty = type_index[i];
        if (ty < 1 || (size_t)ty > ntypes) throw ConfigError("ATOM_TYPE_INDEX out of range");
        const int64_t ico = nb_parm[ntypes * (size_t)(ty - 1) + (size_t)(ty - 1)];
        if (ico <= 0) {
            std::ostringstream m;
            m << "atom " << i << " uses a 10-12 H-bond LJ type (NONBONDED_PARM_INDEX=" << ico
              << "); not supported";
            throw ConfigError(m.str());
        }
        const double a = acoef[(size_t)(ico - 1)], b = bcoef[(size_t)(ico - 1)];
        if (a == 0.0 || b == 0.0) {
            atoms.epsilon.push_back(0.0);
            atoms.sigma.push_back(0.0);
        } else {
            atoms.sigma.push_back(std::pow(a / b, 1.0 / 6.0));
            atoms.epsilon.push_back((b * b) / (4.0 * a));
        }
    }
    for (size_t i = 0; i < natom; i++) {
        AtomMeta m;
        m.element = infer_element(names[i]);
        m.residue_id = residue_id_for_atom(i, residue_pointer);
        m.atom_name = names[i];
        atoms.meta.push_back(m);
    }

    BondedTopology t;
    const auto bonds_h = sec.ints("BONDS_INC_HYDROGEN", 8, 3 * nbonh);
    const auto bonds_a = sec.ints("BONDS_WITHOUT_HYDROGEN", 8, 3 * nbona);
    const auto angles_h = sec.ints("ANGLES_INC_HYDROGEN", 8, 4 * ntheth);
    const auto angles_a = sec.ints("ANGLES_WITHOUT_HYDROGEN", 8, 4 * ntheta);
    const auto dihed_h = sec.ints("DIHEDRALS_INC_HYDROGEN", 8, 5 * nphih);
    const auto dihed_a = sec.ints("DIHEDRALS_WITHOUT_HYDROGEN", 8, 5 * nphia);
    const auto bond_fk = sec.floats("BOND_FORCE_CONSTANT", 16, numbnd);
    const auto bond_eq = sec.floats("BOND_EQUIL_VALUE", 16, numbnd);
    const auto angle_fk = sec.floats("ANGLE_FORCE_CONSTANT", 16, numang);
    const auto angle_eq = sec.floats("ANGLE_EQUIL_VALUE", 16, numang);
    const auto dihed_fk = sec.floats("DIHEDRAL_FORCE_CONSTANT", 16, nptra);
    const auto dihed_per = sec.floats("DIHEDRAL_PERIODICITY", 16, nptra);
    const auto dihed_ph = sec.floats("DIHEDRAL_PHASE", 16, nptra);

    auto type_of = [](int64_t raw, size_t n, const char *what) {
        const int64_t t = raw - 1;
        if (t < 0 || (size_t)t >= n) throw ConfigError(std::string(what) + " type index out of range");
        return (size_t)t;
    };
    // prmtop.rs:138-150: integer division truncates toward zero, `as u32` wraps
    for (const auto *raw : {&bonds_h, &bonds_a})
        for (size_t c = 0; c + 3 <= raw->size(); c += 3) {
            const size_t ty = type_of((*raw)[c + 2], numbnd, "bond");
            t.bond_i.push_back((uint32_t)((*raw)[c] / 3));
            t.bond_j.push_back((uint32_t)((*raw)[c + 1] / 3));
            t.bond_k.push_back(bond_fk[ty]);
            t.bond_r0.push_back(bond_eq[ty]);
        }
    for (const auto *raw : {&angles_h, &angles_a})
        for (size_t c = 0; c + 4 <= raw->size(); c += 4) {
            const size_t ty = type_of((*raw)[c + 3], numang, "angle");
            t.angle_i.push_back((uint32_t)((*raw)[c] / 3));
            t.angle_j.push_back((uint32_t)((*raw)[c + 1] / 3));
            t.angle_k.push_back((uint32_t)((*raw)[c + 2] / 3));
            t.angle_kth.push_back(angle_fk[ty]);
            t.angle_th0.push_back(angle_eq[ty]);
        }
    std::vector<double> scnb(nptra, 2.0), scee(nptra, 1.2);
    if (opt.fidelity) {
        if (sec.has("SCNB_SCALE_FACTOR")) scnb = sec.floats("SCNB_SCALE_FACTOR", 16, nptra);
        if (sec.has("SCEE_SCALE_FACTOR")) scee = sec.floats("SCEE_SCALE_FACTOR", 16, nptra);
    }
    std::map<std::pair<uint32_t, uint32_t>, double> p14;  // (min, max) -> SCNB of the first dihedral
    double scee_seen = 0.0;
    for (const auto *raw : {&dihed_h, &dihed_a})
        for (size_t c = 0; c + 5 <= raw->size(); c += 5) {
            const int64_t *d = raw->data() + c;
            const size_t ty = type_of(d[4], nptra, "dihedral");
            t.dihed_i.push_back((uint32_t)(std::llabs(d[0]) / 3));
            t.dihed_j.push_back((uint32_t)(std::llabs(d[1]) / 3));
            t.dihed_k.push_back((uint32_t)(std::llabs(d[2]) / 3));
            t.dihed_l.push_back((uint32_t)(std::llabs(d[3]) / 3));
            t.dihed_kphi.push_back(dihed_fk[ty]);
            t.dihed_n.push_back(f64_to_u32(std::round(dihed_per[ty])));
            t.dihed_delta.push_back(dihed_ph[ty]);
            if (opt.fidelity && d[2] >= 0 && d[3] >= 0) {
                const uint32_t a = (uint32_t)(std::llabs(d[0]) / 3), b = (uint32_t)(std::llabs(d[3]) / 3);
                const auto key = std::make_pair(std::min(a, b), std::max(a, b));
                if (a != b && !p14.count(key)) p14[key] = scnb[ty];
                scee_seen = scee[ty];
            }
        }
    if (opt.fidelity) {
        double s = 0.0;
        for (const auto &kv : p14) {
            if (!(kv.second > 0.0)) throw ConfigError("SCNB_SCALE_FACTOR must be positive");
            if (s == 0.0) s = kv.second;
            else if (kv.second != s)
                throw ConfigError("1-4 pairs with different SCNB_SCALE_FACTOR values are not supported "
                                  "(the engine applies one 1-4 scale)");
        }
        for (const auto &kv : p14) {
            t.pair14_i.push_back(kv.first.first);
            t.pair14_j.push_back(kv.first.second);
        }
        t.scale_14 = p14.empty() ? 0.0 : 1.0 / s;
        t.scee = scee_seen;
    }
    // prmtop.rs:303-317
    size_t cursor = 0;
    for (size_t i = 0; i < natom; i++) {
        const int64_t cnt = n_excluded[i];
        for (int64_t k = 0; k < cnt; k++) {
            if (cursor >= excluded.size()) throw ConfigError("EXCLUDED_ATOMS_LIST shorter than its counts");
            const int64_t raw = excluded[cursor++];
            if (raw == 0) continue;
            t.excl_i.push_back((uint32_t)i);
            t.excl_j.push_back((uint32_t)(raw - 1));
        }
    }
    return {std::move(atoms), std::move(t)};
}

}  // namespace prmtop

namespace inpcrd {

SimState parse(const std::string &text, size_t expected_n_atoms, bool has_box) {
    const std::vector<std::string> all = lines_of(text);
    if (all.size() < 2)
        throw ConfigError("invalid value for key 'inpcrd': '' (file has fewer than 2 lines "
                          "(missing title or atom-count header))");
    std::istringstream hs(all[1]);
    std::string natom_field;
    hs >> natom_field;
    int64_t natom_i = 0;
    if (!parse_i64(natom_field, natom_i) || natom_i < 0)
        throw ConfigError("invalid value for key 'inpcrd header': '" + natom_field + "' (expected an atom count)");
    const size_t natom = (size_t)natom_i;
    if (natom != expected_n_atoms) {
        std::ostringstream m;
        m << "inpcrd declares " << natom << " atoms but prmtop declares " << expected_n_atoms;
        throw ConfigError(m.str());
    }
    const size_t nval = natom * 3, ncl = (nval + 5) / 6;
    const size_t ndata = all.size() - 2;
    if (ndata < ncl) {
        std::ostringstream m;
        m << "expected " << ncl << " coordinate line(s) for " << natom << " atoms, found " << ndata;
        throw ConfigError(m.str());
    }
    auto block = [&](size_t first, const char *what) {
        std::string joined;
        for (size_t k = 0; k < ncl; k++) joined += all[2 + first + k];
        std::vector<double> v;
        for (const std::string &f : chunk_fields(joined, 12, nval)) {
            double d = 0;
            if (!parse_f64(f, d))
                throw ConfigError(std::string("invalid value for key '") + what + "': '" + f + "' (expected a float)");
            v.push_back(d);
        }
        return v;
    };
    SimState s(natom);
    const std::vector<double> c = block(0, "inpcrd coordinates");
    for (size_t i = 0; i < natom; i++) {
        s.pos_x[i] = c[3 * i];
        s.pos_y[i] = c[3 * i + 1];
        s.pos_z[i] = c[3 * i + 2];
    }
    const size_t remaining = ndata - ncl;
    const size_t n_box = has_box ? 1 : 0;
    const size_t avail = remaining > n_box ? remaining - n_box : 0;
    if (avail >= ncl && ncl > 0) {
        const std::vector<double> v = block(ncl, "inpcrd velocities");
        for (size_t i = 0; i < natom; i++) {
            s.vel_x[i] = v[3 * i] * kAmberVelocityToAngPerPs;
            s.vel_y[i] = v[3 * i + 1] * kAmberVelocityToAngPerPs;
            s.vel_z[i] = v[3 * i + 2] * kAmberVelocityToAngPerPs;
        }
    }
    return s;
}

}  // namespace inpcrd

void promote_hydrogen_bonds(BondedTopology &t, const AtomData &atoms) {
    std::vector<uint32_t> ki, kj;
    std::vector<double> kk, kr;
    for (size_t b = 0; b < t.bond_i.size(); b++) {
        const uint32_t i = t.bond_i[b], j = t.bond_j[b];
        const double r0 = t.bond_r0[b];
        if (i >= atoms.meta.size() || j >= atoms.meta.size()) throw ConfigError("bond index out of range");
        if (atoms.meta[i].element == Element::H || atoms.meta[j].element == Element::H) {
            t.constr_i.push_back(i);
            t.constr_j.push_back(j);
            t.constr_dsq.push_back(r0 * r0);
        } else {
            ki.push_back(i);
            kj.push_back(j);
            kk.push_back(t.bond_k[b]);
            kr.push_back(r0);
        }
    }
    t.bond_i = ki;
    t.bond_j = kj;
    t.bond_k = kk;
    t.bond_r0 = kr;
}

}  // namespace geodesic
