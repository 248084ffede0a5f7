// lattice.hpp - velocity sets and MRT moment bases as compile-time tables.
//
// Direction orderings follow the reference (Items::setc, basic.cu:70-98): D2Q9 in the x-z
// plane (1:+x 2:+z 3:-x 4:-z 5:(+,+) 6:(-,+) 7:(-,-) 8:(+,-)); D3Q27 rest, +x +y -x -y +z -z,
// xy edges, xz edges, yz edges, corners.  D3Q19 and D3Q7 are the first 19 / 7 directions of
// that ordering.  Because every table is constexpr, a fully unrolled loop over directions
// lets the compiler fold the integer moment matrix into adds and subtracts.
#pragma once

#include <type_traits>

#if defined(__HIPCC__)
#define IPL_HD __host__ __device__
#else
#define IPL_HD
#endif

namespace ipl {

struct Dir { int x, y, z; };

constexpr Dir kD2Q9[9] = {{0, 0, 0}, {1, 0, 0}, {0, 0, 1}, {-1, 0, 0}, {0, 0, -1},
                          {1, 0, 1}, {-1, 0, 1}, {-1, 0, -1}, {1, 0, -1}};
constexpr Dir kD3Q27[27] = {
    {0, 0, 0},  {1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {1, -1, 0}, {-1, 1, 0}, {-1, -1, 0},
    {1, 0, 1},  {1, 0, -1}, {-1, 0, 1}, {-1, 0, -1},
    {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1},
    {1, 1, 1},  {1, 1, -1}, {1, -1, 1}, {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};

template <int Q> struct Lattice;

template <> struct Lattice<9> {
    static constexpr int Q = 9, dim = 2, nstress = 2;
    static constexpr IPL_HD Dir c(int k) { return kD2Q9[k]; }
    static constexpr IPL_HD double w(int k) { return k == 0 ? 4.0 / 9.0 : (k < 5 ? 1.0 / 9.0 : 1.0 / 36.0); }
    static constexpr int stress[2] = {7, 8};
};
template <> struct Lattice<27> {
    static constexpr int Q = 27, dim = 3, nstress = 5;
    static constexpr IPL_HD Dir c(int k) { return kD3Q27[k]; }
    static constexpr IPL_HD double w(int k) {
        return k == 0 ? 8.0 / 27.0 : (k < 7 ? 2.0 / 27.0 : (k < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
    }
    static constexpr int stress[5] = {5, 6, 7, 8, 9};
};
template <> struct Lattice<19> {
    static constexpr int Q = 19, dim = 3, nstress = 5;
    static constexpr IPL_HD Dir c(int k) { return kD3Q27[k]; }
    static constexpr IPL_HD double w(int k) { return k == 0 ? 1.0 / 3.0 : (k < 7 ? 1.0 / 18.0 : 1.0 / 36.0); }
    static constexpr int stress[5] = {9, 11, 13, 14, 15};
};
template <> struct Lattice<7> {
    static constexpr int Q = 7, dim = 3, nstress = 0;
    static constexpr IPL_HD Dir c(int k) { return kD3Q27[k]; }
    static constexpr IPL_HD double w(int k) { return k == 0 ? 0.25 : 0.125; }
    static constexpr int stress[1] = {-1};
};

template <int Q> constexpr IPL_HD int comp(int k, int axis) {
    return axis == 0 ? Lattice<Q>::c(k).x : (axis == 1 ? Lattice<Q>::c(k).y : Lattice<Q>::c(k).z);
}

// opposite direction (set_k_inv, basic.cu:112-121)
template <int Q> constexpr IPL_HD int opposite(int k) {
    const Dir a = Lattice<Q>::c(k);
    for (int j = 0; j < Q; ++j) {
        const Dir b = Lattice<Q>::c(j);
        if (b.x == -a.x && b.y == -a.y && b.z == -a.z) return j;
    }
    return 0;
}

// direction with the given components, -1 if absent
template <int Q> constexpr IPL_HD int find_dir(int x, int y, int z) {
    for (int j = 0; j < Q; ++j) {
        const Dir b = Lattice<Q>::c(j);
        if (b.x == x && b.y == y && b.z == z) return j;
    }
    return -1;
}

// direction k with its component along `axis` reversed
template <int Q> constexpr IPL_HD int mirror(int k, int axis) {
    Dir a = Lattice<Q>::c(k);
    if (axis == 0) a.x = -a.x; else if (axis == 1) a.y = -a.y; else a.z = -a.z;
    return find_dir<Q>(a.x, a.y, a.z);
}

// unit direction along axis with sign s
template <int Q> constexpr IPL_HD int axis_dir(int axis, int s) {
    return axis == 0 ? find_dir<Q>(s, 0, 0) : (axis == 1 ? find_dir<Q>(0, s, 0) : find_dir<Q>(0, 0, s));
}

// compile-time loop: f(std::integral_constant<int, I>) for I = 0 .. N-1.  Table lookups at the
// indices it passes (moment<Q>(r, k) and the like) fold to constants: a plain loop that the
// compiler does not unroll fully (27 x 27 terms) would evaluate them at run time.
template <int N, int I = 0, typename F>
IPL_HD inline void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// MRT moment basis, row i evaluated at direction k.  D2Q9 and D3Q27 reproduce the rows of
// the reference's input/MRT_D3Q9.csv and MRT_D3Q27.csv; D3Q19 is the d'Humieres et al.
// (2002) basis in the same ordering.
template <int Q> constexpr IPL_HD double moment(int i, int k) {
    const Dir d = Lattice<Q>::c(k);
    const double x = d.x, y = d.y, z = d.z, c2 = x * x + y * y + z * z;
    if (Q == 9) {
        switch (i) {
        case 0: return 1; case 1: return -4 + 3 * c2; case 2: return 4 - 10.5 * c2 + 4.5 * c2 * c2;
        case 3: return x; case 4: return (-5 + 3 * c2) * x; case 5: return z; case 6: return (-5 + 3 * c2) * z;
        case 7: return x * x - z * z; default: return x * z;
        }
    } else if (Q == 27) {
        switch (i) {
        case 0: return 1; case 1: return x; case 2: return y; case 3: return z; case 4: return c2 - 2;
        case 5: return 3 * x * x - c2; case 6: return y * y - z * z; case 7: return x * y; case 8: return y * z;
        case 9: return x * z; case 10: return x * (3 * c2 - 7); case 11: return y * (3 * c2 - 7);
        case 12: return z * (3 * c2 - 7);
        case 13: return x * (9 * c2 * c2 - 39 * c2 + 38) / 2; case 14: return y * (9 * c2 * c2 - 39 * c2 + 38) / 2;
        case 15: return z * (9 * c2 * c2 - 39 * c2 + 38) / 2;
        case 16: return (3 * c2 * c2 - 11 * c2 + 8) / 2;
        case 17: return (9 * c2 * c2 * c2 - 45 * c2 * c2 + 60 * c2 - 16) / 2;
        case 18: return (3 * x * x - c2) * (3 * c2 - 5); case 19: return (y * y - z * z) * (3 * c2 - 5);
        case 20: return x * y * (3 * c2 - 8); case 21: return y * z * (3 * c2 - 8); case 22: return x * z * (3 * c2 - 8);
        case 23: return x * (y * y - z * z); case 24: return y * (z * z - x * x); case 25: return z * (x * x - y * y);
        default: return x * y * z;
        }
    } else {  // 19
        switch (i) {
        case 0: return 1; case 1: return 19 * c2 - 30; case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
        case 3: return x; case 4: return (5 * c2 - 9) * x; case 5: return y; case 6: return (5 * c2 - 9) * y;
        case 7: return z; case 8: return (5 * c2 - 9) * z;
        case 9: return 3 * x * x - c2; case 10: return (3 * c2 - 5) * (3 * x * x - c2);
        case 11: return y * y - z * z; case 12: return (3 * c2 - 5) * (y * y - z * z);
        case 13: return x * y; case 14: return y * z; case 15: return x * z;
        case 16: return x * (y * y - z * z); case 17: return y * (z * z - x * x); default: return z * (x * x - y * y);
        }
    }
}

template <int Q> constexpr IPL_HD double moment_norm2(int i) {
    double s = 0;
    for (int k = 0; k < Q; ++k) s += moment<Q>(i, k) * moment<Q>(i, k);
    return s;
}

// M^-1[k][i] = M[i][k] / |row i|^2 (the rows are orthogonal)
template <int Q> constexpr IPL_HD double moment_inv(int k, int i) { return moment<Q>(i, k) / moment_norm2<Q>(i); }

template <int Q> constexpr IPL_HD bool is_stress(int i) {
    for (int s = 0; s < Lattice<Q>::nstress; ++s)
        if (Lattice<Q>::stress[s] == i) return true;
    return false;
}

// run-time Q forms for host code
inline double moment_row_norm2(int Q, int i) {
    return Q == 9 ? moment_norm2<9>(i) : (Q == 19 ? moment_norm2<19>(i) : moment_norm2<27>(i));
}
inline bool moment_row_is_stress(int Q, int i) {
    return Q == 9 ? is_stress<9>(i) : (Q == 19 ? is_stress<19>(i) : is_stress<27>(i));
}

}  // namespace ipl
