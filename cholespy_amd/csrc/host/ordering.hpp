// Fill-reducing ordering in the role CHOLMOD's NESDIS plays for the reference
// (src/cholesky_solver.cpp:206-209: simplicial LL', one method, CHOLMOD_NESDIS).
// CHOLMOD/METIS are not available, so this is an independent nested
// dissection built from published methods:
//   * pseudo-peripheral start vertices and rooted level structures
//     (Gibbs-Poole-Stockmeyer 1976; George & Liu, "Computer Solution of Large
//     Sparse Positive Definite Systems", 1981, ch. 4 and 8),
//   * level-set vertex separators trimmed to a minimal separator
//     (George 1973, "Nested dissection of a regular finite element mesh"),
//   * exact minimum degree on small leaves (Tinney-Walker scheme 2).
// The resulting permutation differs from CHOLMOD's; the solve phase only needs
// some valid symmetric permutation, and every consumer takes it as input.
#pragma once
#include <vector>

namespace cholespy_amd {

// Symmetric adjacency structure without self loops (compressed rows).
struct Graph {
    int n = 0;
    std::vector<int> xadj;   // n + 1
    std::vector<int> adj;
};

struct NDOptions {
    int leaf_size = 48;         // stop dissecting at this many vertices
    double min_balance = 0.25;  // each side of a separator keeps >= this fraction
};

// Returns perm with perm[k] = original vertex placed at position k.
std::vector<int> nested_dissection(const Graph &g, const NDOptions &opt = NDOptions());

}  // namespace cholespy_amd
