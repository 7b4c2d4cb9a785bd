// Drop-in C++ host API with the names and signatures of the reference's
// X-TED_GPU/TED_C++.h, implemented over the xted C ABI (include/xted.h) on
// one MI355X. The CUDA/Thrust declarations of the reference header (device
// kernels, queues, streams) have no counterpart: scheduling and kernels live
// behind the C ABI. Like the reference header this one brings namespace std
// into scope, so code written against it compiles unchanged.
//
// Behaviour kept from the reference:
//   * standard_ted_1 clears the caller's label vectors (X-TED_GPU/TED.cu:93-94).
//   * standard_ted returns n for an empty x tree and m for an empty y tree
//     (X-TED_GPU/TED.cu:119-124).
//   * parallel_standard_ted_test prints the reference's two timing lines
//     (X-TED_GPU/parallel_X_TED.cu:224,392); set XTED_QUIET=1 to silence them.
// Behaviour changed on purpose:
//   * the distance and the full tree-distance table are returned (the
//     reference GPU path returns a zero matrix, X-TED_GPU/parallel_X_TED.cu:396-397);
//   * inputs must be preorder-numbered trees (the format's own contract,
//     README.md:42); anything else throws std::invalid_argument, as do
//     malformed adjacency lines (the reference's stoi throws there too).
#ifndef XTED_TED_CPP_H
#define XTED_TED_CPP_H

#include <string>
#include <vector>

using namespace std;

// Device time of the last call's forest-DP kernels, in ms
// (X-TED_GPU/TED.cu:3, accumulated at X-TED_GPU/parallel_X_TED.cu:429-443).
extern float total_milliseconds;

// X-TED_GPU/TED_C++.h:24-28
int min3(int a, int b, int c);
vector<int> outermost_right_leaves(vector<vector<int>>& adj);
vector<int> key_roots(vector<int>& orl);
vector<vector<int>> standard_ted_1(vector<string>& x_node, vector<vector<int>>& x_adj,
                                   vector<string>& y_node, vector<vector<int>>& y_adj,
                                   int num_threads, int parallel_version);
int standard_ted(vector<string>& x_node, vector<vector<int>>& x_adj, vector<string>& y_node,
                 vector<vector<int>>& y_adj, int num_threads, int parallel_version);

// X-TED_GPU/TED_C++.h:50. Delta must be a label-equality matrix
// (Delta[i][j] == 0 exactly when x_i and y_j carry the same label).
vector<vector<int>> parallel_standard_ted_test(vector<int>& x_orl, vector<int>& x_kr,
                                               vector<int>& y_orl, vector<int>& y_kr,
                                               vector<vector<int>>& Delta,
                                               vector<vector<int>>& D,
                                               vector<vector<int>>& D_tree, int m, int n,
                                               int num_threads, vector<vector<int>>& x_adj,
                                               vector<vector<int>>& y_adj);

// X-TED_GPU/TED_C++.h:53-57 (dataset text formats)
int ToInt(const string& str);
vector<int> To_Vector(string str);
vector<string> splitting(string str);
vector<vector<int>> pre_process(string str);
vector<string> node_process(string str);

// X-TED_GPU/TED_C++.h:60: computes pair (0, 1) of the two files and prints
// "The distance = d" (the reference's CPU driver line, X-TED_CPU/TED_test.cpp:45).
void test(char* input_1, char* input_2);

#endif  // XTED_TED_CPP_H
