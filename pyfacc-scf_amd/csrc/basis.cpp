// Gaussian-94 basis reader and shell construction.
//
// Semantics follow the reference's build_basis (basis.f90:9-92,
// read_shell :95-130, read_sp_shell :133-190):
//   * element blocks are visited in file order; a block whose label matches
//     no atom of the molecule is skipped;
//   * every shell line of a block creates one shell per atom of that
//     element, atoms in input order;
//   * an SP line creates the S shells of all those atoms, then their P shells,
//     sharing exponents;
//   * the scale factor on the shell line is read and ignored (basis.f90:75).
// Basis functions follow shell order, Cartesian components x, y, z for p
// (cartesian.f90:7-23).  Exponents written with a Fortran 'D' are accepted.
// Unlike the reference, a skipped block is skipped up to its '****', so an
// "S"/"P" shell line of a foreign element is never mistaken for a sulfur or
// phosphorus block header; shells of angular momentum above p are rejected
// with an error instead of being misread.
#include <dlfcn.h>
#include <sys/stat.h>

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <tuple>

#include "pf_common.h"

namespace pf {

namespace {

bool file_exists(const std::string &p) {
    struct stat st;
    return stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

std::string lib_dir() {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&resolve_basis_path), &info) && info.dli_fname) {
        std::string f = info.dli_fname;
        auto pos = f.find_last_of('/');
        if (pos != std::string::npos) return f.substr(0, pos);
    }
    return ".";
}

double parse_real(std::string s) {
    for (auto &c : s)
        if (c == 'D' || c == 'd') c = 'E';
    char *end = nullptr;
    double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str()) throw Error("basis file: cannot parse number '" + s + "'");
    return v;
}

std::vector<std::string> split(const std::string &line) {
    std::istringstream is(line);
    std::vector<std::string> out;
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

std::string trim(const std::string &s) {
    size_t b = s.find_first_not_of(" \t\r\n"), e = s.find_last_not_of(" \t\r\n");
    return b == std::string::npos ? std::string() : s.substr(b, e - b + 1);
}

double dfact(int n) {
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

// Primitive norm of the (l,0,0) Cartesian component (gaussian.f90:80-98).
double prim_norm(double a, int l) {
    return std::pow(2.0 * a / M_PI, 0.75) * std::pow(4.0 * a, 0.5 * l) / std::sqrt(dfact(2 * l - 1));
}

}  // namespace

std::string resolve_basis_path(const std::string &name) {
    std::vector<std::string> cands;
    cands.push_back("./data/basis_set/" + name + ".txt");  // main.f90:33 (CWD-relative)
    if (const char *env = std::getenv("PYFACC_DATA")) cands.push_back(std::string(env) + "/basis_set/" + name + ".txt");
    cands.push_back(lib_dir() + "/../data/basis_set/" + name + ".txt");
    cands.push_back(name);
    for (auto &c : cands)
        if (file_exists(c)) return c;
    std::string msg = "basis set '" + name + "' not found; tried:";
    for (auto &c : cands) msg += " " + c;
    throw Error(msg);
}

std::vector<std::string> read_basis_file(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw Error("cannot open basis file " + path);
    std::vector<std::string> lines;
    for (std::string l; std::getline(in, l);) lines.push_back(l);
    return lines;
}

// Parsed basis files, cached per path and keyed by size and modification
// time (an edited file is read again): parsing the 1100-line STO-3G file
// costs about as much as an SCF iteration of (H2O)64.
std::shared_ptr<const BasisBlocks> basis_blocks_cached(const std::string &path) {
    struct stat st;
    if (stat(path.c_str(), &st) != 0) throw Error("cannot open basis file " + path);
    const long long size = st.st_size, mtime = (long long)st.st_mtim.tv_sec * 1000000000LL + st.st_mtim.tv_nsec;
    static std::mutex mu;
    static std::map<std::string, std::tuple<long long, long long, std::shared_ptr<const BasisBlocks>>> cache;
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(path);
    if (it != cache.end() && std::get<0>(it->second) == size && std::get<1>(it->second) == mtime)
        return std::get<2>(it->second);
    auto blocks = std::make_shared<const BasisBlocks>(parse_basis_blocks(read_basis_file(path), path));
    cache[path] = std::make_tuple(size, mtime, blocks);
    return blocks;
}

Basis build_basis(const std::string &path, const std::vector<std::string> &species, const Molecule &mol,
                  bool renormalise) {
    return build_basis_blocks(*basis_blocks_cached(path), species, mol, renormalise);
}

BasisBlocks parse_basis_blocks(const std::vector<std::string> &lines, const std::string &path) {
    BasisBlocks blocks;
    blocks.path = path;
    size_t i = 0;
    while (i < lines.size()) {
        std::string line = lines[i++];
        if (trim(line).empty() || line[0] == '!') continue;
        auto tok = split(line);
        if (tok.empty() || tok[0] == "****") continue;
        BasisBlocks::Block blk;
        blk.label = tok[0];
        // walk the block to its terminator
        while (i < lines.size()) {
            std::string sl = lines[i++];
            if (sl.find("****") != std::string::npos) break;
            auto st = split(sl);
            if (st.size() < 2) continue;
            BasisBlocks::ShellLine sh;
            sh.kind = st[0];
            sh.np = std::atoi(st[1].c_str());
            if (sh.np <= 0 || sh.np > KMAX) throw Error("basis file: bad primitive count on line '" + sl + "'");
            for (int k = 0; k < sh.np; ++k) {
                if (i >= lines.size()) throw Error("basis file: truncated shell for " + blk.label);
                std::vector<double> r;
                for (auto &t : split(lines[i++])) r.push_back(parse_real(t));
                sh.rows.push_back(r);
            }
            blk.shells.push_back(sh);
        }
        blocks.blocks.push_back(blk);
    }
    return blocks;
}

Basis build_basis_lines(const std::vector<std::string> &lines, const std::string &path,
                        const std::vector<std::string> &species, const Molecule &mol, bool renormalise) {
    return build_basis_blocks(parse_basis_blocks(lines, path), species, mol, renormalise);
}

Basis build_basis_blocks(const BasisBlocks &blocks, const std::vector<std::string> &species, const Molecule &mol,
                         bool renormalise) {
    Basis basis;
    for (const auto &blk : blocks.blocks) {
        std::vector<int> atoms;
        for (int a = 0; a < (int)species.size(); ++a)
            if (species[a] == blk.label) atoms.push_back(a);
        if (atoms.empty()) continue;
        for (const auto &sh : blk.shells) {
            auto add = [&](int l, int col) {
                for (int a : atoms) {
                    Shell s;
                    s.l = l;
                    s.np = sh.np;
                    s.atom = a;
                    for (int d = 0; d < 3; ++d) s.xyz[d] = mol.xyz[3 * a + d];
                    for (int k = 0; k < sh.np; ++k) {
                        if ((int)sh.rows[k].size() <= col)
                            throw Error("basis file: missing coefficient for " + blk.label);
                        s.exp[k] = sh.rows[k][0];
                        s.coef[k] = sh.rows[k][col];
                    }
                    basis.shells.push_back(s);
                }
            };
            if (sh.kind == "S") add(0, 1);
            else if (sh.kind == "P") add(1, 1);
            else if (sh.kind == "SP") { add(0, 1); add(1, 2); }
            else if (sh.kind == "D") add(2, 1);  // basis.f90:76-82 has no D branch; the Cartesian code is general
            else
                throw Error("basis file: shell type '" + sh.kind + "' of " + blk.label +
                            " is not supported (s, p and d shells only)");
        }
    }
    // every atom must have received functions
    std::vector<int> seen(species.size(), 0);
    for (auto &s : basis.shells) seen[s.atom] = 1;
    for (size_t a = 0; a < species.size(); ++a)
        if (!seen[a]) throw Error("basis file " + blocks.path + " has no functions for element '" + species[a] + "'");

    int bf = 0;
    for (auto &s : basis.shells) {
        s.bf = bf;
        bf += ncart(s.l);
        for (int k = 0; k < s.np; ++k) s.coef[k] *= prim_norm(s.exp[k], s.l);
        if (renormalise) {
            double ss = 0.0;
            for (int p = 0; p < s.np; ++p)
                for (int q = 0; q < s.np; ++q) {
                    double z = s.exp[p] + s.exp[q];
                    ss += s.coef[p] * s.coef[q] * std::pow(M_PI / z, 1.5) * dfact(2 * s.l - 1) / std::pow(2.0 * z, s.l);
                }
            double sc = 1.0 / std::sqrt(ss);
            for (int k = 0; k < s.np; ++k) s.coef[k] *= sc;
        }
    }
    basis.nbf = bf;
    return basis;
}

double nuclear_repulsion(const Molecule &mol) {
    double e = 0.0;
    for (int i = 0; i < mol.natoms; ++i)
        for (int j = i + 1; j < mol.natoms; ++j) {
            double d2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                double d = mol.xyz[3 * i + k] - mol.xyz[3 * j + k];
                d2 += d * d;
            }
            double r = std::sqrt(d2);
            if (r > 1e-10) e += double(mol.Z[i]) * mol.Z[j] / r;
        }
    return e;
}

}  // namespace pf
