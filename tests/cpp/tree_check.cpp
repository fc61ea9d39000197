// Drives the C++ class of include/hpdmk_tree.hpp the way the reference's
// shift-energy test drives HPDMKPtTree (test/hpdmk_test.cpp:374-465): build,
// form expansions, total energy and its parts, then a chain of trial moves,
// each evaluated with eval_shift_energy and applied with update_shift, and the
// energy of the final state.
//
// usage: tree_check <input file> double|float
// input (text): n L digits n_per_leaf, then n lines "x y z q", then m and
// m lines "i dx dy dz".  Output lines: "E e w d r", one "dE v" per move, "E_end e".
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "hpdmk_tree.hpp"

template <typename Real>
int run(std::ifstream& in) {
    int n, digits, leaf;
    double L;
    in >> n >> L >> digits >> leaf;
    std::vector<Real> r(3 * (size_t)n), q(n);
    for (int i = 0; i < n; ++i) {
        double x, y, z, c;
        in >> x >> y >> z >> c;
        r[3 * i] = (Real)x;
        r[3 * i + 1] = (Real)y;
        r[3 * i + 2] = (Real)z;
        q[i] = (Real)c;
    }
    if (!in) throw std::runtime_error("malformed input (header or particle lines)");
    HPDMKParams params = hpdmk_default_params();
    params.n_per_leaf = leaf;
    params.digits = digits;
    params.L = L;
    params.init = DIRECT;
    hpdmk_amd::HPDMKPtTree<Real> tree(params, r, q);
    tree.form_outgoing_pw();
    tree.form_incoming_pw();
    std::printf("E %.17g %.17g %.17g %.17g\n", (double)tree.eval_energy(), (double)tree.eval_energy_window(),
                (double)tree.eval_energy_diff(), (double)tree.eval_energy_res());
    int m;
    in >> m;
    for (int k = 0; k < m; ++k) {
        long i;
        double dx, dy, dz;
        in >> i >> dx >> dy >> dz;
        if (!in) throw std::runtime_error("malformed move line");
        const Real dE = tree.eval_shift_energy(i, (Real)dx, (Real)dy, (Real)dz);
        tree.update_shift(i, (Real)dx, (Real)dy, (Real)dz);
        std::printf("dE %.17g\n", (double)dE);
    }
    std::printf("E_end %.17g\n", (double)tree.eval_energy());
    // the reference throws on a bad index; so does this class
    try {
        tree.eval_shift_energy(n, 0, 0, 0);
        std::printf("bad_index_not_rejected\n");
    } catch (const std::runtime_error&) {
        std::printf("bad_index_rejected\n");
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input> double|float\n", argv[0]);
        return 2;
    }
    std::ifstream in(argv[1]);
    if (!in) {
        std::fprintf(stderr, "cannot open %s\n", argv[1]);
        return 2;
    }
    try {
        return std::strcmp(argv[2], "float") == 0 ? run<float>(in) : run<double>(in);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
}
