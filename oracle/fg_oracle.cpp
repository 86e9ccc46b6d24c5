// fg_oracle.cpp -- TEST INFRASTRUCTURE ONLY (see fg_oracle.h).
//
// A plain f64 restatement of the reference's algorithm, written from the reference's
// source read as text. Each routine cites the file:line it follows. The product library
// never links or calls this file.
//
// Loops over independent blocks (variable blocks, edges, cameras, points) run on OpenMP
// threads. Every sum is still taken term by term in the order of the sequential loop it
// replaces (a block's terms in emission order), so results are bit-identical for any
// thread count; orc_set_threads(1) is the single-threaded restatement.
//
// Third-party arithmetic that the reference delegates (faer 0.24.4 CG/LLT/triangular
// solves, nalgebra 0.34.2 quaternion rotation) is restated from the published
// algorithms: textbook PCG (Saad, Iterative Methods, Alg. 9.1, x0 = 0, stop when
// ||r||2 <= tol*||b||2), column Cholesky, forward/back substitution, and the
// unit-quaternion rotation matrix of exp(w/2).
#include "fg_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

namespace {

const double kNaN = std::numeric_limits<double>::quiet_NaN();
const size_t kNone = std::numeric_limits<size_t>::max();

bool is_finite(double x) { return std::isfinite(x); }

// src/optimization.rs:653-662
int norm_inf(const double* v, size_t n, double* out) {
  double norm = 0.0;
  for (size_t i = 0; i < n; ++i) {
    if (!is_finite(v[i])) return ORC_INVALID_EVALUATION;
    norm = std::max(norm, std::fabs(v[i]));
  }
  if (out) *out = norm;
  return ORC_OK;
}

// src/linear_model.rs:325-337 (hypot chain)
double norm_l2(const double* v, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s = std::hypot(s, v[i]);
  return s;
}
double relative(double norm, double reference) { return reference == 0.0 ? norm : norm / reference; }

// ---------------------------------------------------------------------------------------
// Projection: examples/bal/model.rs:116-184
// ---------------------------------------------------------------------------------------
void skew(const double v[3], double k[9]) {
  k[0] = 0; k[1] = -v[2]; k[2] = v[1];
  k[3] = v[2]; k[4] = 0; k[5] = -v[0];
  k[6] = -v[1]; k[7] = v[0]; k[8] = 0;
}
void matmul3(const double a[9], const double b[9], double c[9]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += a[3 * i + k] * b[3 * k + j];
      c[3 * i + j] = s;
    }
}

// Unit quaternion exp(w/2) -> rotation matrix (nalgebra UnitQuaternion::from_scaled_axis
// followed by to_rotation_matrix, examples/bal/model.rs:141-143).
void rotation_from_scaled_axis(const double w[3], double r[9]) {
  double v[3] = {w[0] / 2, w[1] / 2, w[2] / 2};
  double nn = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  double eps = std::numeric_limits<double>::epsilon();
  double qi = 0, qj = 0, qk = 0, qw = 1;
  if (!(nn <= eps * eps)) {
    double n = std::sqrt(nn);
    double s = std::sin(n) / n;
    qi = v[0] * s; qj = v[1] * s; qk = v[2] * s; qw = std::cos(n);
  }
  double ww = qw * qw, ii = qi * qi, jj = qj * qj, kk = qk * qk;
  double ij = qi * qj * 2, wk = qw * qk * 2, wj = qw * qj * 2, ik = qi * qk * 2, jk = qj * qk * 2,
         wi = qw * qi * 2;
  r[0] = ww + ii - jj - kk; r[1] = ij - wk; r[2] = wj + ik;
  r[3] = wk + ij; r[4] = ww - ii + jj - kk; r[5] = jk - wi;
  r[6] = ik - wj; r[7] = wi + jk; r[8] = ww - ii - jj + kk;
}

struct Projection {
  double cam[9];
  double rot[9];
  double right[9];
  // examples/bal/model.rs:126-146
  explicit Projection(const double c[9]) {
    std::memcpy(cam, c, sizeof(cam));
    const double* w = c;
    double t2 = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    double a, b;
    if (t2 < 1e-8) {
      a = 0.5 - t2 / 24. + t2 * t2 / 720.;
      b = 1. / 6. - t2 / 120. + t2 * t2 / 5040.;
    } else {
      double t = std::sqrt(t2);
      a = (1. - std::cos(t)) / t2;
      b = (t - std::sin(t)) / (t2 * t);
    }
    double k[9], kk[9];
    skew(w, k);
    matmul3(k, k, kk);
    for (int i = 0; i < 9; ++i) right[i] = (i % 4 == 0 ? 1.0 : 0.0) - a * k[i] + b * kk[i];
    rotation_from_scaled_axis(w, rot);
  }
  void camera_point(const double x[3], double p[3]) const {
    for (int i = 0; i < 3; ++i)
      p[i] = rot[3 * i] * x[0] + rot[3 * i + 1] * x[1] + rot[3 * i + 2] * x[2] + cam[3 + i];
  }
  // examples/bal/model.rs:147-156
  int value(const double x[3], double pixel[2]) const {
    double p[3];
    camera_point(x, p);
    double q0 = -p[0] / p[2], q1 = -p[1] / p[2];
    double r2 = q0 * q0 + q1 * q1;
    double s = cam[6] * (1. + cam[7] * r2 + cam[8] * r2 * r2);
    pixel[0] = s * q0;
    pixel[1] = s * q1;
    if (p[2] == 0. || !is_finite(pixel[0]) || !is_finite(pixel[1])) return ORC_INVALID_EVALUATION;
    return ORC_OK;
  }
  // examples/bal/model.rs:157-184; jc is 2x9 and jp 2x3, both column-major.
  void jacobians(const double x[3], double jc[18], double jp[6]) const {
    double p[3];
    camera_point(x, p);
    double q0 = -p[0] / p[2], q1 = -p[1] / p[2];
    double r2 = q0 * q0 + q1 * q1;
    double f = cam[6];
    double radial = 1. + cam[7] * r2 + cam[8] * r2 * r2;
    double c2 = 2. * cam[7] + 4. * cam[8] * r2;
    double dq[4] = {f * (radial + c2 * q0 * q0), f * (c2 * q0 * q1), f * (c2 * q1 * q0),
                    f * (radial + c2 * q1 * q1)};  // row-major 2x2
    double m[6] = {-1. / p[2], 0., p[0] / (p[2] * p[2]), 0., -1. / p[2], p[1] / (p[2] * p[2])};
    double dp[6];  // row-major 2x3
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) dp[3 * i + j] = dq[2 * i] * m[j] + dq[2 * i + 1] * m[3 + j];
    // dp * -R * skew(X) * Jr, evaluated left to right.
    double a[6], b[6], c[6];
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += dp[3 * i + k] * (-rot[3 * k + j]);
        a[3 * i + j] = s;
      }
    double sk[9];
    skew(x, sk);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += a[3 * i + k] * sk[3 * k + j];
        b[3 * i + j] = s;
      }
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0;
        for (int k = 0; k < 3; ++k) s += b[3 * i + k] * right[3 * k + j];
        c[3 * i + j] = s;
      }
    for (int j = 0; j < 3; ++j) {
      jc[2 * j] = c[j];
      jc[2 * j + 1] = c[3 + j];
      jc[2 * (3 + j)] = dp[j];
      jc[2 * (3 + j) + 1] = dp[3 + j];
    }
    jc[12] = radial * q0; jc[13] = radial * q1;
    jc[14] = f * r2 * q0; jc[15] = f * r2 * q1;
    jc[16] = f * r2 * r2 * q0; jc[17] = f * r2 * r2 * q1;
    for (int j = 0; j < 3; ++j) {
      double s0 = 0, s1 = 0;
      for (int k = 0; k < 3; ++k) {
        s0 += dp[k] * rot[3 * k + j];
        s1 += dp[3 + k] * rot[3 * k + j];
      }
      jp[2 * j] = s0;
      jp[2 * j + 1] = s1;
    }
  }
};

// ---------------------------------------------------------------------------------------
// Linear model: src/linear_model.rs
// ---------------------------------------------------------------------------------------
struct Block {
  size_t row, col, rows, cols, start;
};
// Block index for the threaded loops: runs of blocks that share their first row (one
// emission each), and the blocks of every column start in block order. `clean` holds when
// the column groups are disjoint and each has one width, i.e. every column is covered by the
// blocks of exactly one group; only then do per-group loops visit a column's terms in the
// sequential order.
struct BlockIndex {
  bool built = false, clean = false;
  std::vector<size_t> row_runs;             // first block of each run, then blocks.size()
  std::vector<size_t> group_col, group_width;
  std::vector<size_t> group_off, group_blocks;  // CSR over groups, block order
  std::vector<size_t> group_of_col;         // column start -> group (kNone elsewhere)
};

// Stable counting sort of items by key: CSR offsets (keys + 1) and items in input order.
void bucket(size_t keys, size_t n, const size_t* key, std::vector<size_t>& off, std::vector<size_t>& items) {
  off.assign(keys + 1, 0);
  for (size_t i = 0; i < n; ++i)
    if (key[i] != kNone) off[key[i] + 1]++;
  for (size_t k = 0; k < keys; ++k) off[k + 1] += off[k];
  items.assign(off[keys], 0);
  std::vector<size_t> fill(off.begin(), off.end() - 1);
  for (size_t i = 0; i < n; ++i)
    if (key[i] != kNone) items[fill[key[i]]++] = i;
}

struct Jacobian {
  size_t rows = 0, cols = 0;
  std::vector<Block> blocks;
  std::vector<double> values;
  mutable BlockIndex index;

  void invalidate() { index.built = false; }
  const BlockIndex& indexed() const {
    BlockIndex& ix = index;
    if (ix.built) return ix;
    size_t nb = blocks.size();
    ix.row_runs.clear();
    for (size_t b = 0; b < nb; ++b)
      if (b == 0 || blocks[b].row != blocks[b - 1].row) ix.row_runs.push_back(b);
    ix.row_runs.push_back(nb);
    std::vector<size_t> start_of(cols, kNone);
    std::vector<size_t> key(nb, kNone);
    ix.group_col.clear();
    ix.group_width.clear();
    ix.clean = true;
    for (size_t b = 0; b < nb; ++b) {
      if (blocks[b].cols == 0) continue;
      size_t c = blocks[b].col;
      if (c >= cols) { ix.clean = false; continue; }
      if (start_of[c] == kNone) {
        start_of[c] = ix.group_col.size();
        ix.group_col.push_back(c);
        ix.group_width.push_back(blocks[b].cols);
      } else if (ix.group_width[start_of[c]] != blocks[b].cols) {
        ix.clean = false;
      }
      key[b] = start_of[c];
    }
    bucket(ix.group_col.size(), nb, key.data(), ix.group_off, ix.group_blocks);
    size_t end = 0;
    for (size_t c = 0; c < cols && ix.clean; ++c) {
      if (start_of[c] == kNone) continue;
      if (c < end || c + ix.group_width[start_of[c]] > cols) ix.clean = false;
      end = c + ix.group_width[start_of[c]];
    }
    ix.group_of_col.swap(start_of);
    ix.built = true;
    return ix;
  }

  // src/linear_model.rs:39-55
  void apply(const double* x, double* y) const {
    std::fill(y, y + rows, 0.0);
    const BlockIndex& ix = indexed();
    size_t runs = ix.row_runs.size() - 1;
    // rows of different runs are disjoint (each run is one emission)
#pragma omp parallel for schedule(static, 1024)
    for (size_t r = 0; r < runs; ++r)
      for (size_t bi = ix.row_runs[r]; bi < ix.row_runs[r + 1]; ++bi) {
        const Block& b = blocks[bi];
        for (size_t j = 0; j < b.cols; ++j) {
          double v = x[b.col + j];
          for (size_t i = 0; i < b.rows; ++i) y[b.row + i] += values[b.start + j * b.rows + i] * v;
        }
      }
  }
  // src/linear_model.rs:64-87
  void transpose_apply(const double* y, double* x) const {
    std::fill(x, x + cols, 0.0);
    const BlockIndex& ix = indexed();
    if (!ix.clean) {
      for (const Block& b : blocks)
        for (size_t j = 0; j < b.cols; ++j) {
          double s = 0;
          for (size_t i = 0; i < b.rows; ++i) s += values[b.start + j * b.rows + i] * y[b.row + i];
          x[b.col + j] += s;
        }
      return;
    }
    size_t groups = ix.group_col.size();
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t g = 0; g < groups; ++g)
      for (size_t k = ix.group_off[g]; k < ix.group_off[g + 1]; ++k) {
        const Block& b = blocks[ix.group_blocks[k]];
        for (size_t j = 0; j < b.cols; ++j) {
          double s = 0;
          for (size_t i = 0; i < b.rows; ++i) s += values[b.start + j * b.rows + i] * y[b.row + i];
          x[b.col + j] += s;
        }
      }
  }
};

struct StepWorkspace {  // src/linear_model.rs:285-318
  std::vector<double> step, product, normal_residual, reference;
  void prepare_verification(size_t rows, size_t cols) {
    product.assign(rows, 0.0);
    normal_residual.assign(cols, 0.0);
    reference.assign(cols, 0.0);
  }
  void scale_residual(const std::vector<double>& scales) {
    for (size_t j = 0; j < normal_residual.size(); ++j) {
      normal_residual[j] /= scales[j];
      reference[j] /= scales[j];
    }
  }
};
struct StepVerification {
  double predicted_reduction;
  bool finite_step;
  double residual_norm, reference_norm;
};

struct LinearizedModel {
  Jacobian jacobian;
  std::vector<double> negative_residual, gradient, scales;
  void prepare(size_t dim) {  // :169-172
    jacobian.cols = dim;
    gradient.assign(dim, 0.0);
  }
  void clear() {  // :173-179
    jacobian.rows = 0;
    jacobian.blocks.clear();
    jacobian.values.clear();
    jacobian.invalidate();
    negative_residual.clear();
    std::fill(gradient.begin(), gradient.end(), 0.0);
  }
  // :180-210; values of block i are column-major rows x widths[i].
  void accumulate(const double* residual, size_t rows, size_t nblocks, const size_t* cols,
                  const size_t* widths, const double* values) {
    size_t row = jacobian.rows;
    jacobian.rows = row + rows;
    jacobian.invalidate();
    for (size_t i = 0; i < rows; ++i) negative_residual.push_back(-residual[i]);
    size_t offset = 0;
    for (size_t bi = 0; bi < nblocks; ++bi) {
      jacobian.blocks.push_back(Block{row, cols[bi], rows, widths[bi], jacobian.values.size()});
      for (size_t j = 0; j < widths[bi]; ++j) {
        double g = 0;
        for (size_t i = 0; i < rows; ++i) {
          double v = values[offset + j * rows + i];
          jacobian.values.push_back(v);
          g += v * residual[i];
        }
        gradient[cols[bi] + j] += g;
      }
      offset += rows * widths[bi];
    }
  }
  int prepare_scales(double floor) {  // :211-230
    if (!is_finite(floor) || floor <= 0.0) return ORC_INVALID_OPTIONS;
    scales.assign(jacobian.cols, 0.0);
    const BlockIndex& ix = jacobian.indexed();
    if (ix.clean) {
      // a column's chain runs over its group's blocks in block order, as below
      size_t groups = ix.group_col.size();
#pragma omp parallel for schedule(dynamic, 64)
      for (size_t g = 0; g < groups; ++g)
        for (size_t k = ix.group_off[g]; k < ix.group_off[g + 1]; ++k) {
          const Block& b = jacobian.blocks[ix.group_blocks[k]];
          for (size_t c = 0; c < b.cols; ++c) {
            double& n = scales[b.col + c];
            for (size_t r = 0; r < b.rows; ++r) n = std::hypot(n, jacobian.values[b.start + c * b.rows + r]);
          }
        }
    } else {
      for (const Block& b : jacobian.blocks)
        for (size_t c = 0; c < b.cols; ++c) {
          double& n = scales[b.col + c];
          for (size_t r = 0; r < b.rows; ++r) n = std::hypot(n, jacobian.values[b.start + c * b.rows + r]);
        }
    }
    if (norm_inf(scales.data(), scales.size(), nullptr) != ORC_OK) return ORC_INVALID_EVALUATION;
    for (double& s : scales) s = std::max(s, floor);
    return ORC_OK;
  }
  double predicted_reduction(const double* step, double* product) const {  // :233-250
    jacobian.apply(step, product);
    double pred = 0;
    for (size_t j = 0; j < jacobian.cols; ++j) pred -= gradient[j] * step[j];
    for (size_t i = 0; i < jacobian.rows; ++i) pred -= (0.5 * product[i]) * product[i];
    return pred;
  }
  StepVerification verify_step(StepWorkspace& w, double lambda) const {  // :254-282
    size_t n = jacobian.cols;
    StepVerification v;
    v.finite_step = norm_inf(w.step.data(), n, nullptr) == ORC_OK;
    v.predicted_reduction = predicted_reduction(w.step.data(), w.product.data());
    for (size_t i = 0; i < jacobian.rows; ++i) w.product[i] -= negative_residual[i];
    jacobian.transpose_apply(w.product.data(), w.normal_residual.data());
    double rn = 0, gn = 0;
    for (size_t j = 0; j < n; ++j) {
      double s = scales[j];
      w.normal_residual[j] += (lambda * (s * w.step[j])) * s;
      rn = std::hypot(rn, w.normal_residual[j]);
      gn = std::hypot(gn, gradient[j]);
    }
    w.reference = gradient;
    v.residual_norm = rn;
    v.reference_norm = gn;
    return v;
  }
};

// ---------------------------------------------------------------------------------------
// Block factors: src/linear_model/block_factors.rs
// ---------------------------------------------------------------------------------------
// Column Cholesky of a column-major w x w matrix, lower triangle in place. Returns false on
// a nonpositive or nonfinite pivot (the failure classes faer's LLT reports).
bool cholesky_lower(double* a, size_t w) {
  for (size_t j = 0; j < w; ++j) {
    double d = a[j * w + j];
    for (size_t k = 0; k < j; ++k) d -= a[k * w + j] * a[k * w + j];
    if (!(d > 0.0) || !is_finite(d)) return false;
    d = std::sqrt(d);
    a[j * w + j] = d;
    for (size_t i = j + 1; i < w; ++i) {
      double s = a[j * w + i];
      for (size_t k = 0; k < j; ++k) s -= a[k * w + i] * a[k * w + j];
      a[j * w + i] = s / d;
    }
  }
  return true;
}
// L y = b (forward) and L^T y = b (backward), L column-major lower.
void solve_lower(const double* l, size_t w, double* b) {
  for (size_t i = 0; i < w; ++i) {
    double s = b[i];
    for (size_t k = 0; k < i; ++k) s -= l[k * w + i] * b[k];
    b[i] = s / l[i * w + i];
  }
}
void solve_upper_transpose(const double* l, size_t w, double* b) {
  for (size_t ii = w; ii-- > 0;) {
    double s = b[ii];
    for (size_t k = ii + 1; k < w; ++k) s -= l[ii * w + k] * b[k];
    b[ii] = s / l[ii * w + ii];
  }
}

struct VariableBlock {
  size_t col, width, start;
  bool fallback;
};
struct BlockFactors {
  std::vector<VariableBlock> blocks;
  std::vector<size_t> owners;
  std::vector<double> gram, factors;
  int prepare(const Jacobian& jac, const std::vector<double>& scales) {  // :47-111
    blocks.clear();
    gram.clear();
    owners.assign(jac.cols, kNone);
    std::vector<size_t> ids(jac.blocks.size(), kNone);
    for (size_t bi = 0; bi < jac.blocks.size(); ++bi) {
      const Block& b = jac.blocks[bi];
      if (b.cols == 0) continue;
      size_t end = b.col + b.cols;
      if (end > jac.cols) return ORC_DIMENSION_MISMATCH;
      size_t owner = owners[b.col];
      size_t id;
      if (owner == kNone) {
        for (size_t c = b.col; c < end; ++c)
          if (owners[c] != kNone) return ORC_DIMENSION_MISMATCH;
        id = blocks.size();
        size_t start = gram.size();
        gram.resize(start + b.cols * b.cols, 0.0);
        blocks.push_back(VariableBlock{b.col, b.cols, start, false});
        for (size_t c = b.col; c < end; ++c) owners[c] = id;
      } else {
        if (blocks[owner].col != b.col || blocks[owner].width != b.cols) return ORC_DIMENSION_MISMATCH;
        id = owner;
      }
      ids[bi] = id;
    }
    auto add = [&](const Block& b, size_t start) {
      for (size_t col = 0; col < b.cols; ++col)
        for (size_t row = col; row < b.cols; ++row) {
          double s = 0;
          for (size_t k = 0; k < b.rows; ++k)
            s += (jac.values[b.start + col * b.rows + k] / scales[b.col + col]) *
                 (jac.values[b.start + row * b.rows + k] / scales[b.col + row]);
          gram[start + col * b.cols + row] += s;
        }
    };
    const BlockIndex& ix = jac.indexed();
    if (ix.clean) {
      // the blocks of a variable block are exactly its column group, visited in block order
      size_t nv = blocks.size();
#pragma omp parallel for schedule(dynamic, 64)
      for (size_t v = 0; v < nv; ++v) {
        size_t g = ix.group_of_col[blocks[v].col];
        for (size_t k = ix.group_off[g]; k < ix.group_off[g + 1]; ++k)
          add(jac.blocks[ix.group_blocks[k]], blocks[v].start);
      }
    } else {
      for (size_t bi = 0; bi < jac.blocks.size(); ++bi)
        if (ids[bi] != kNone) add(jac.blocks[bi], blocks[ids[bi]].start);
    }
    factors.assign(gram.size(), 0.0);
    return ORC_OK;
  }
  size_t factor(double lambda) {  // :113-138
    factors = gram;
    size_t fallbacks = 0;
    size_t nv = blocks.size();
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : fallbacks)
    for (size_t bi = 0; bi < nv; ++bi) {
      VariableBlock& b = blocks[bi];
      double* v = factors.data() + b.start;
      for (size_t j = 0; j < b.width; ++j) v[j * b.width + j] += lambda;
      bool ok = norm_inf(v, b.width * b.width, nullptr) == ORC_OK && cholesky_lower(v, b.width);
      b.fallback = !ok;
      if (!b.fallback)
        for (size_t j = 0; j < b.width; ++j)
          if (!is_finite(v[j * b.width + j]) || v[j * b.width + j] <= 0.0) b.fallback = true;
      fallbacks += b.fallback ? 1 : 0;
    }
    return fallbacks;
  }
  void apply_factor_inverse(double* rhs, bool transpose) const {  // :142-160
    size_t nv = blocks.size();
#pragma omp parallel for schedule(dynamic, 256)
    for (size_t bi = 0; bi < nv; ++bi) {
      const VariableBlock& b = blocks[bi];
      if (b.fallback) continue;
      const double* l = factors.data() + b.start;
      if (transpose) solve_lower(l, b.width, rhs + b.col);
      else solve_upper_transpose(l, b.width, rhs + b.col);
    }
  }
  // :178-204 over coordinates [offset, offset+n)
  void solve_normal(double* rhs, size_t offset, size_t n, double lambda) const {
    size_t nv = blocks.size();
#pragma omp parallel for schedule(dynamic, 256)
    for (size_t bi = 0; bi < nv; ++bi) {
      const VariableBlock& b = blocks[bi];
      if (b.col < offset || b.col >= offset + n) continue;
      const double* l = factors.data() + b.start;
      double* part = rhs + (b.col - offset);
      solve_lower(l, b.width, part);
      solve_upper_transpose(l, b.width, part);
    }
    if (lambda != 0.0) {
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < n; ++i)
        if (owners[offset + i] == kNone) rhs[i] /= lambda;
    }
  }
};

// ---------------------------------------------------------------------------------------
// Schur backend: src/schur.rs
// ---------------------------------------------------------------------------------------
struct Edge {
  size_t retained, eliminated, retained_jacobian, eliminated_jacobian;
};

}  // namespace

struct orc_schur {
  size_t max_iterations = 1000;
  double relative_tolerance = 1e-10;  // src/schur.rs:155-157 for f64
  size_t retained, rw, ew;
  LinearizedModel model;
  BlockFactors blocks;
  StepWorkspace work;
  std::vector<Edge> edges;
  // edges of each retained / eliminated block in edge order (built by prepare_damping)
  std::vector<size_t> retained_off, retained_edges, eliminated_off, eliminated_edges;
  std::vector<double> cross, reduced_rhs, reduced_step, scaled_work, reduced_residual;
  bool ready = false;
  size_t solves = 0, iterations = 0;
  orc_schur_diagnostics diag;

  void reset_diag(double damping) {
    diag.termination = ORC_CG_NONE;
    diag.iterations = 0;
    diag.damping = damping;
    diag.estimated_relative_residual = kNaN;
    diag.reduced_relative_residual = kNaN;
    diag.scaled_relative_residual = kNaN;
    diag.original_relative_residual = kNaN;
  }

  void index_edges() {
    size_t ne = edges.size(), dim = model.jacobian.cols;
    std::vector<size_t> key(ne);
    for (size_t e = 0; e < ne; ++e) key[e] = edges[e].retained / rw;
    bucket(retained / rw, ne, key.data(), retained_off, retained_edges);
    for (size_t e = 0; e < ne; ++e) key[e] = edges[e].eliminated / ew;
    bucket((dim - retained) / ew, ne, key.data(), eliminated_off, eliminated_edges);
  }
  // src/schur.rs:558-582. The sequential loop visits edges in order; an output block's terms
  // come from its own edges only, so each block sums its edges in edge order on one thread.
  void cross_apply(double* out, const double* rhs, bool transpose, double alpha) const {
    const std::vector<size_t>& off = transpose ? eliminated_off : retained_off;
    const std::vector<size_t>& list = transpose ? eliminated_edges : retained_edges;
    size_t nblocks = off.empty() ? 0 : off.size() - 1;
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t blk = 0; blk < nblocks; ++blk)
      for (size_t k = off[blk]; k < off[blk + 1]; ++k) {
        size_t e = list[k];
        for (size_t p = 0; p < ew; ++p)
          for (size_t c = 0; c < rw; ++c) {
            double value = alpha * cross[(e * ew + p) * rw + c];
            if (transpose) out[edges[e].eliminated + p] += value * rhs[edges[e].retained + c];
            else out[edges[e].retained + c] += value * rhs[edges[e].eliminated + p];
          }
      }
  }
  // Reduced::apply, src/schur.rs:606-652
  void reduced_apply(double lambda, const double* x, double* out) const {
    size_t n = model.jacobian.cols, elim = n - retained;
    for (size_t i = 0; i < retained; ++i) out[i] = lambda * x[i];
    size_t nv = blocks.blocks.size();
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t bi = 0; bi < nv; ++bi) {
      const VariableBlock& b = blocks.blocks[bi];
      if (b.col >= retained) continue;
      const double* g = blocks.gram.data() + b.start;
      for (size_t j = 0; j < b.width; ++j)
        for (size_t i = 0; i < b.width; ++i) {
          double value = g[std::min(i, j) * b.width + std::max(i, j)];
          out[b.col + i] += value * x[b.col + j];
        }
    }
    std::vector<double> tmp(elim, 0.0);
    cross_apply(tmp.data(), x, true, 1.0);
    blocks.solve_normal(tmp.data(), retained, elim, lambda);
    cross_apply(out, tmp.data(), false, -1.0);
  }
  // RetainedPreconditioner::apply, src/schur.rs:670-677
  void precondition(double lambda, const double* r, double* z) const {
    std::copy(r, r + retained, z);
    blocks.solve_normal(z, 0, retained, lambda);
  }

  int factor_and_reduce_rhs(double lambda) {  // :204-247
    size_t n = model.jacobian.cols;
    if (blocks.factor(lambda) != 0) {
      diag.termination = ORC_CG_NOT_POSITIVE_DEFINITE;
      return ORC_LINEAR_SOLVE_FAILED;
    }
    for (size_t j = 0; j < n; ++j) scaled_work[j] = -model.gradient[j] / model.scales[j];
    blocks.solve_normal(scaled_work.data() + retained, retained, n - retained, lambda);
    std::copy(scaled_work.begin(), scaled_work.begin() + retained, reduced_rhs.begin());
    cross_apply(reduced_rhs.data(), scaled_work.data() + retained, false, -1.0);
    return ORC_OK;
  }

  // Textbook PCG in place of faer's conjugate_gradient (src/schur.rs:249-310).
  void solve_reduced(double lambda) {
    size_t n = retained;
    std::fill(reduced_step.begin(), reduced_step.end(), 0.0);
    std::vector<double> r(n), z(n), p(n), q(n);
    reduced_apply(lambda, reduced_step.data(), q.data());
    for (size_t i = 0; i < n; ++i) r[i] = reduced_rhs[i] - q[i];
    double bnorm = norm_l2(reduced_rhs.data(), n);
    double tol = relative_tolerance * bnorm;
    double rnorm = norm_l2(r.data(), n);
    if (rnorm <= tol) {
      diag.termination = ORC_CG_CONVERGED;
      diag.estimated_relative_residual = relative(rnorm, bnorm);
      return;
    }
    precondition(lambda, r.data(), z.data());
    p = z;
    double rz = 0;
    for (size_t i = 0; i < n; ++i) rz += r[i] * z[i];
    for (size_t it = 0; it < max_iterations; ++it) {
      reduced_apply(lambda, p.data(), q.data());
      double pq = 0;
      for (size_t i = 0; i < n; ++i) pq += p[i] * q[i];
      if (!(pq > 0.0) || !is_finite(pq)) {
        diag.termination = ORC_CG_NOT_POSITIVE_DEFINITE;
        return;
      }
      double alpha = rz / pq;
      for (size_t i = 0; i < n; ++i) {
        reduced_step[i] += alpha * p[i];
        r[i] -= alpha * q[i];
      }
      ++iterations;
      ++diag.iterations;
      rnorm = norm_l2(r.data(), n);
      if (rnorm <= tol) {
        diag.termination = ORC_CG_CONVERGED;
        diag.estimated_relative_residual = relative(rnorm, bnorm);
        return;
      }
      precondition(lambda, r.data(), z.data());
      double rz_new = 0;
      for (size_t i = 0; i < n; ++i) rz_new += r[i] * z[i];
      double beta = rz_new / rz;
      rz = rz_new;
      for (size_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    diag.termination = ORC_CG_ITERATION_LIMIT;
  }

  void recover_step(double lambda) {  // :312-344
    size_t n = model.jacobian.cols;
    std::copy(reduced_step.begin(), reduced_step.end(), scaled_work.begin());
    for (size_t j = retained; j < n; ++j) scaled_work[j] = -model.gradient[j] / model.scales[j];
    // Edge offsets of eliminated blocks are relative to the retained prefix.
    cross_apply(scaled_work.data() + retained, reduced_step.data(), true, -1.0);
    blocks.solve_normal(scaled_work.data() + retained, retained, n - retained, lambda);
    for (size_t j = 0; j < n; ++j) work.step[j] = scaled_work[j] / model.scales[j];
  }

  int verify_step(double lambda, double* predicted) {  // :346-397
    reduced_apply(lambda, reduced_step.data(), reduced_residual.data());
    for (size_t i = 0; i < retained; ++i) reduced_residual[i] -= reduced_rhs[i];
    diag.reduced_relative_residual =
        relative(norm_l2(reduced_residual.data(), retained), norm_l2(reduced_rhs.data(), retained));
    StepVerification v = model.verify_step(work, lambda);
    diag.original_relative_residual = relative(v.residual_norm, v.reference_norm);
    work.scale_residual(model.scales);
    diag.scaled_relative_residual =
        relative(norm_l2(work.normal_residual.data(), work.normal_residual.size()),
                 norm_l2(work.reference.data(), work.reference.size()));
    bool bad = !v.finite_step || !is_finite(v.predicted_reduction);
    for (double r : {diag.reduced_relative_residual, diag.scaled_relative_residual,
                     diag.original_relative_residual})
      if (!is_finite(r)) bad = true;
    if (bad) {
      diag.termination = ORC_CG_NON_FINITE;
      return ORC_LINEAR_SOLVE_FAILED;
    }
    if (diag.termination != ORC_CG_CONVERGED) return ORC_LINEAR_SOLVE_FAILED;
    *predicted = v.predicted_reduction;
    return ORC_OK;
  }
};

extern "C" {

int orc_projection(const double cam[9], double rotation[9], double right[9]) {
  Projection p(cam);
  std::memcpy(rotation, p.rot, sizeof(p.rot));
  std::memcpy(right, p.right, sizeof(p.right));
  return ORC_OK;
}
int orc_project(const double cam[9], const double point[3], double pixel[2]) {
  return Projection(cam).value(point, pixel);
}
void orc_jacobians(const double cam[9], const double point[3], double jc[18], double jp[6]) {
  Projection(cam).jacobians(point, jc, jp);
}

orc_schur* orc_schur_new(size_t retained, size_t rw, size_t ew) {
  orc_schur* s = new orc_schur();
  s->retained = retained;
  s->rw = rw;
  s->ew = ew;
  s->reset_diag(0.0);
  return s;
}
void orc_schur_free(orc_schur* s) { delete s; }
void orc_schur_set_options(orc_schur* s, size_t max_iterations, double tol) {
  s->max_iterations = max_iterations;
  s->relative_tolerance = tol;
}
int orc_schur_prepare(orc_schur* s, size_t dim) {  // src/schur.rs:409-431
  s->ready = false;
  s->solves = s->iterations = 0;
  s->reset_diag(0.0);
  if (s->rw == 0 || s->ew == 0 || s->retained > dim || s->retained % s->rw != 0 ||
      (dim - s->retained) % s->ew != 0 || s->max_iterations == 0 ||
      !(s->relative_tolerance >= 0.0 && s->relative_tolerance < 1.0))
    return ORC_INVALID_OPTIONS;
  s->model.prepare(dim);
  s->work.step.assign(dim, 0.0);
  s->reduced_rhs.assign(s->retained, 0.0);
  s->reduced_step.assign(s->retained, 0.0);
  s->reduced_residual.assign(s->retained, 0.0);
  s->scaled_work.assign(dim, 0.0);
  return ORC_OK;
}
void orc_schur_clear(orc_schur* s) {  // :432-437
  s->ready = false;
  s->edges.clear();
  s->cross.clear();
  s->model.clear();
}
int orc_schur_accumulate(orc_schur* s, const double* residual, size_t rows, size_t nblocks,
                         const size_t* cols, const size_t* widths, const double* values) {
  // src/schur.rs:438-482
  s->ready = false;
  size_t left_col = kNone, left_idx = kNone, right_col = kNone, right_idx = kNone;
  size_t start = s->model.jacobian.blocks.size();
  for (size_t i = 0; i < nblocks; ++i) {
    size_t col = cols[i], end = col + widths[i];
    if (end > s->model.jacobian.cols) return ORC_DIMENSION_MISMATCH;
    if (col < s->retained) {
      if (col % s->rw != 0 || widths[i] != s->rw || end > s->retained || left_col != kNone)
        return ORC_UNSUPPORTED_STRUCTURE;
      left_col = col;
      left_idx = start + i;
    } else {
      if ((col - s->retained) % s->ew != 0 || widths[i] != s->ew || right_col != kNone)
        return ORC_UNSUPPORTED_STRUCTURE;
      right_col = col - s->retained;
      right_idx = start + i;
    }
  }
  s->model.accumulate(residual, rows, nblocks, cols, widths, values);
  if (left_col != kNone && right_col != kNone)
    s->edges.push_back(Edge{left_col, right_col, left_idx, right_idx});
  return ORC_OK;
}
int orc_schur_gradient_norm(const orc_schur* s, double* out) {
  return norm_inf(s->model.gradient.data(), s->model.gradient.size(), out);
}
int orc_schur_prepare_damping(orc_schur* s, double floor) {  // :490-526
  s->ready = false;
  int st = s->model.prepare_scales(floor);
  if (st != ORC_OK) return st;
  s->work.prepare_verification(s->model.jacobian.rows, s->model.jacobian.cols);
  st = s->blocks.prepare(s->model.jacobian, s->model.scales);
  if (st != ORC_OK) return st;
  s->cross.assign(s->rw * s->ew * s->edges.size(), 0.0);
  s->index_edges();
  const LinearizedModel& m = s->model;
  size_t ne = s->edges.size();
#pragma omp parallel for schedule(static, 1024)
  for (size_t e = 0; e < ne; ++e) {
    const Block& left = m.jacobian.blocks[s->edges[e].retained_jacobian];
    const Block& right = m.jacobian.blocks[s->edges[e].eliminated_jacobian];
    for (size_t p = 0; p < s->ew; ++p)
      for (size_t c = 0; c < s->rw; ++c) {
        double sum = 0;
        for (size_t row = 0; row < left.rows; ++row)
          sum += (m.jacobian.values[left.start + c * left.rows + row] / m.scales[left.col + c]) *
                 (m.jacobian.values[right.start + p * right.rows + row] / m.scales[right.col + p]);
        s->cross[(e * s->ew + p) * s->rw + c] = sum;
      }
  }
  if (norm_inf(s->cross.data(), s->cross.size(), nullptr) != ORC_OK) return ORC_INVALID_EVALUATION;
  s->ready = true;
  return ORC_OK;
}
int orc_schur_solve_damped(orc_schur* s, double lambda, double* delta, double* predicted) {
  // src/schur.rs:192-202, 527-545
  s->reset_diag(lambda);
  if (!s->ready || !is_finite(lambda) || lambda < 0.0) return ORC_INVALID_OPTIONS;
  s->solves++;
  int st = s->factor_and_reduce_rhs(lambda);
  if (st != ORC_OK) return st;
  s->solve_reduced(lambda);
  s->recover_step(lambda);
  double pred = 0;
  st = s->verify_step(lambda, &pred);
  if (st != ORC_OK) return st;
  if (delta) std::copy(s->work.step.begin(), s->work.step.end(), delta);
  if (predicted) *predicted = pred;
  return ORC_OK;
}
int orc_schur_solve(orc_schur* s, double* delta) {  // :486-489
  int st = orc_schur_prepare_damping(s, 1e-6);
  if (st != ORC_OK) return st;
  return orc_schur_solve_damped(s, 0.0, delta, nullptr);
}
void orc_schur_diagnostics_get(const orc_schur* s, orc_schur_diagnostics* d) { *d = s->diag; }
void orc_schur_statistics(const orc_schur* s, size_t* solves, size_t* iterations) {
  *solves = s->solves;
  *iterations = s->iterations;
}
size_t orc_schur_dimension(const orc_schur* s) { return s->model.jacobian.cols; }
size_t orc_schur_rows(const orc_schur* s) { return s->model.jacobian.rows; }
void orc_schur_gradient(const orc_schur* s, double* out) {
  std::copy(s->model.gradient.begin(), s->model.gradient.end(), out);
}
void orc_schur_scales(const orc_schur* s, double* out) {
  std::copy(s->model.scales.begin(), s->model.scales.end(), out);
}
void orc_schur_negative_residual(const orc_schur* s, double* out) {
  std::copy(s->model.negative_residual.begin(), s->model.negative_residual.end(), out);
}
int orc_schur_reduced_apply(const orc_schur* s, double lambda, const double* x, double* y) {
  if (!s->ready) return ORC_INVALID_OPTIONS;
  // solve_normal needs factors at this lambda: factor a copy.
  orc_schur* m = const_cast<orc_schur*>(s);
  std::vector<double> saved = m->blocks.factors;
  std::vector<VariableBlock> saved_blocks = m->blocks.blocks;
  if (m->blocks.factor(lambda) != 0) {
    m->blocks.factors = saved;
    m->blocks.blocks = saved_blocks;
    return ORC_LINEAR_SOLVE_FAILED;
  }
  s->reduced_apply(lambda, x, y);
  m->blocks.factors = saved;
  m->blocks.blocks = saved_blocks;
  return ORC_OK;
}
void orc_schur_reduced_rhs(const orc_schur* s, double* out) {
  std::copy(s->reduced_rhs.begin(), s->reduced_rhs.end(), out);
}
int orc_schur_gram_block(const orc_schur* s, size_t col, double* out, size_t* width) {
  for (const VariableBlock& b : s->blocks.blocks)
    if (b.col == col) {
      std::copy(s->blocks.gram.begin() + b.start, s->blocks.gram.begin() + b.start + b.width * b.width, out);
      *width = b.width;
      return ORC_OK;
    }
  *width = 0;
  return ORC_DIMENSION_MISMATCH;
}
void orc_schur_jacobian_apply(const orc_schur* s, const double* x, double* y) { s->model.jacobian.apply(x, y); }
void orc_schur_jacobian_transpose_apply(const orc_schur* s, const double* y, double* x) {
  s->model.jacobian.transpose_apply(y, x);
}
// src/lsmr.rs:147-174
void orc_schur_damped_apply(const orc_schur* s, double lambda, int scaled, const double* x, double* y) {
  const LinearizedModel& m = s->model;
  size_t n = m.jacobian.cols, rows = m.jacobian.rows;
  double root = std::sqrt(lambda);
  if (scaled) {
    std::vector<double> xs(n);
    for (size_t j = 0; j < n; ++j) xs[j] = x[j] / m.scales[j];
    m.jacobian.apply(xs.data(), y);
  } else {
    m.jacobian.apply(x, y);
  }
  for (size_t j = 0; j < n; ++j) y[rows + j] = (scaled ? root : root * m.scales[j]) * x[j];
}
// src/lsmr.rs:183-205
void orc_schur_damped_transpose_apply(const orc_schur* s, double lambda, int scaled, const double* y,
                                      double* x) {
  const LinearizedModel& m = s->model;
  size_t n = m.jacobian.cols, rows = m.jacobian.rows;
  double root = std::sqrt(lambda);
  m.jacobian.transpose_apply(y, x);
  for (size_t j = 0; j < n; ++j) {
    if (scaled) x[j] /= m.scales[j];
    x[j] += (scaled ? root : root * m.scales[j]) * y[rows + j];
  }
}
size_t orc_schur_factor(orc_schur* s, double lambda) { return s->blocks.factor(lambda); }
void orc_schur_apply_factor_inverse(const orc_schur* s, double* rhs, int transpose) {
  s->blocks.apply_factor_inverse(rhs, transpose != 0);
}
int orc_schur_prepare_blocks(orc_schur* s, const double* scales) {
  s->model.scales.assign(scales, scales + s->model.jacobian.cols);
  return s->blocks.prepare(s->model.jacobian, s->model.scales);
}
void orc_schur_gram_set(orc_schur* s, size_t index, double value) { s->blocks.gram[index] = value; }
int orc_schur_prepare_scales(orc_schur* s, double floor) { return s->model.prepare_scales(floor); }
double orc_schur_verify(orc_schur* s, const double* step, double lambda, double* normal_residual,
                        double* reference, double* residual_norm, double* reference_norm) {
  size_t n = s->model.jacobian.cols;
  s->work.step.assign(step, step + n);
  s->work.prepare_verification(s->model.jacobian.rows, n);
  StepVerification v = s->model.verify_step(s->work, lambda);
  std::copy(s->work.normal_residual.begin(), s->work.normal_residual.end(), normal_residual);
  std::copy(s->work.reference.begin(), s->work.reference.end(), reference);
  *residual_norm = v.residual_norm;
  *reference_norm = v.reference_norm;
  return v.predicted_reduction;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// LSMR backend over the same linearized model: src/lsmr.rs:215-280 (solve), :430-468
// (recover_and_verify), :560-595 (solve_damped), :535-552 (undamped solve). The iteration is
// faer 0.24.4's `matrix_free::lsmr` (a dependency absent from the reference tree), restated
// from its published algorithm: Fong & Saunders, "LSMR: an iterative algorithm for sparse
// least-squares problems", SIAM J. Sci. Comput. 33(5), 2011, Algorithm LSMR without local
// reorthogonalization, x0 = 0, right preconditioner P (solver coordinates y, x = P y), stopped
// when the normal-residual estimate |zeta_bar| <= tol * ||(A P)^T b||, which is the relative
// estimate the reference reports (src/lsmr.rs:82-85).
// ---------------------------------------------------------------------------------------
namespace {
double nrm2(const std::vector<double>& v) {
  double s = 0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}
// mode: 0 identity [J; sqrt(l) D], 1 diagonal [J D^-1; sqrt(l) I], 2 block (1 plus P = L^-T);
// undamped: operator J alone, rhs of length rows.
struct LsmrOperator {
  const orc_schur* s;
  int mode;
  bool undamped;
  double lambda;
  size_t n() const { return s->model.jacobian.cols; }
  size_t m() const { return s->model.jacobian.rows + (undamped ? 0 : n()); }
  void precond(double* x, bool transpose) const {
    if (mode == 2) s->blocks.apply_factor_inverse(x, transpose);
  }
  // y = A P x
  void apply(const double* x, double* y) const {
    std::vector<double> px(x, x + n());
    precond(px.data(), false);
    if (undamped) s->model.jacobian.apply(px.data(), y);
    else orc_schur_damped_apply(s, lambda, mode != 0, px.data(), y);
  }
  // x = P^T A^T y
  void transpose_apply(const double* y, double* x) const {
    if (undamped) s->model.jacobian.transpose_apply(y, x);
    else orc_schur_damped_transpose_apply(s, lambda, mode != 0, y, x);
    precond(x, true);
  }
};
struct LsmrResult {
  bool converged;
  size_t iterations;
  double abs_residual, rel_residual;
};
LsmrResult lsmr(const LsmrOperator& op, const double* b, double* x_out, double tol, size_t max_iter) {
  const size_t m = op.m(), n = op.n();
  std::vector<double> u(b, b + m), v(n), y(n, 0.0), h(n), hbar(n, 0.0), tmp_m(m), tmp_n(n);
  double beta = nrm2(u);
  if (beta > 0) for (double& e : u) e /= beta;
  op.transpose_apply(u.data(), v.data());
  double alpha = nrm2(v);
  if (alpha > 0) for (double& e : v) e /= alpha;
  double alphabar = alpha, zetabar = alpha * beta, rho = 1, rhobar = 1, cbar = 1, sbar = 0;
  h = v;
  const double norm_atb = alpha * beta;
  LsmrResult r{false, 0, std::fabs(zetabar), 0.0};
  if (!(norm_atb > 0)) {  // A^T b = 0: the zero step is exact
    r.converged = true;
    r.abs_residual = 0;
    std::fill(x_out, x_out + n, 0.0);
    return r;
  }
  for (size_t k = 1; k <= max_iter; ++k) {
    op.apply(v.data(), tmp_m.data());
    for (size_t i = 0; i < m; ++i) u[i] = tmp_m[i] - alpha * u[i];
    beta = nrm2(u);
    if (beta > 0) for (double& e : u) e /= beta;
    op.transpose_apply(u.data(), tmp_n.data());
    for (size_t j = 0; j < n; ++j) v[j] = tmp_n[j] - beta * v[j];
    alpha = nrm2(v);
    if (alpha > 0) for (double& e : v) e /= alpha;
    const double rho_old = rho, rhobar_old = rhobar;
    rho = std::hypot(alphabar, beta);
    const double c = alphabar / rho, sn = beta / rho;
    const double theta_new = sn * alpha;
    alphabar = c * alpha;
    const double thetabar = sbar * rho;
    rhobar = std::hypot(cbar * rho, theta_new);
    cbar = cbar * rho / rhobar;
    sbar = theta_new / rhobar;
    const double zeta = cbar * zetabar;
    zetabar = -sbar * zetabar;
    const double f1 = thetabar * rho / (rho_old * rhobar_old), f2 = zeta / (rho * rhobar), f3 = theta_new / rho;
    for (size_t j = 0; j < n; ++j) {
      hbar[j] = h[j] - f1 * hbar[j];
      y[j] += f2 * hbar[j];
      h[j] = v[j] - f3 * h[j];
    }
    r.iterations = k;
    r.abs_residual = std::fabs(zetabar);
    r.rel_residual = r.abs_residual / norm_atb;
    if (r.abs_residual <= tol * norm_atb) {
      r.converged = true;
      break;
    }
  }
  op.precond(y.data(), false);
  std::copy(y.begin(), y.end(), x_out);
  return r;
}
}  // namespace

extern "C" {

// src/lsmr.rs:560-595 with recover_and_verify (:430-468); mode as LsmrOperator.
int orc_lsmr_solve_damped(orc_schur* s, double lambda, int mode, size_t max_iterations, double tol,
                          double* delta, double* predicted, orc_lsmr_diagnostics* d) {
  *d = orc_lsmr_diagnostics{ORC_LSMR_NONE, 0, 0, kNaN, kNaN, kNaN, kNaN, kNaN, kNaN, kNaN, kNaN, kNaN, 0};
  if (mode < 0 || mode > 2 || max_iterations == 0 || !(tol >= 0.0 && tol < 1.0)) return ORC_INVALID_OPTIONS;
  if (!std::isfinite(lambda) || lambda < 0.0) return ORC_INVALID_OPTIONS;
  const LinearizedModel& m = s->model;
  const size_t n = m.jacobian.cols, rows = m.jacobian.rows;
  d->damping = lambda;
  d->min_scale = *std::min_element(m.scales.begin(), m.scales.end());
  d->max_scale = *std::max_element(m.scales.begin(), m.scales.end());
  const double root = std::sqrt(lambda);
  for (size_t j = 0; j < n; ++j)
    if (!std::isfinite(mode != 0 ? root : root * m.scales[j])) {
      d->termination = ORC_LSMR_NON_FINITE;
      return ORC_LINEAR_SOLVE_FAILED;
    }
  if (mode == 2) d->block_fallbacks = s->blocks.factor(lambda);
  std::vector<double> rhs(rows + n, 0.0);
  std::copy(m.negative_residual.begin(), m.negative_residual.end(), rhs.begin());
  LsmrOperator op{s, mode, false, lambda};
  std::vector<double> step(n, 0.0);
  bool ok_solve = true;
  if (rows + n > 0 && n > 0) {
    s->solves += 1;
    LsmrResult r = lsmr(op, rhs.data(), step.data(), tol, max_iterations);
    s->iterations += r.iterations;
    d->iterations = r.iterations;
    d->estimated_absolute_residual = r.abs_residual;
    d->estimated_relative_residual = r.rel_residual;
    d->finite_step = norm_inf(step.data(), n, nullptr) == ORC_OK;
    const bool finite = d->finite_step && std::isfinite(r.abs_residual) && std::isfinite(r.rel_residual);
    d->termination = finite ? (r.converged ? ORC_LSMR_CONVERGED : ORC_LSMR_ITERATION_LIMIT) : ORC_LSMR_NON_FINITE;
    ok_solve = d->termination == ORC_LSMR_CONVERGED;
  } else {
    d->termination = ORC_LSMR_CONVERGED;
    d->finite_step = 1;
  }
  // recover_and_verify
  if (mode != 0)
    for (size_t j = 0; j < n; ++j) step[j] /= m.scales[j];
  s->work.step = step;
  s->work.prepare_verification(rows, n);
  StepVerification v = m.verify_step(s->work, lambda);
  d->finite_step = v.finite_step;
  if (mode != 0) s->work.scale_residual(m.scales);
  if (mode == 2) {
    s->blocks.apply_factor_inverse(s->work.normal_residual.data(), true);
    s->blocks.apply_factor_inverse(s->work.reference.data(), true);
  }
  const double scaled_norm = nrm2(s->work.normal_residual), scaled_reference = nrm2(s->work.reference);
  auto relative = [](double a, double b) { return b == 0.0 ? a : a / b; };  // linear_model.rs relative()
  d->original_absolute_residual = v.residual_norm;
  d->original_relative_residual = relative(v.residual_norm, v.reference_norm);
  d->scaled_absolute_residual = scaled_norm;
  d->scaled_relative_residual = relative(scaled_norm, scaled_reference);
  std::copy(step.begin(), step.end(), delta);
  if (!v.finite_step || !std::isfinite(v.residual_norm) || !std::isfinite(scaled_norm)) {
    d->termination = ORC_LSMR_NON_FINITE;
    return ORC_LINEAR_SOLVE_FAILED;
  }
  if (predicted) *predicted = v.predicted_reduction;
  return ok_solve ? ORC_OK : ORC_LINEAR_SOLVE_FAILED;
}

// src/lsmr.rs:535-552: identity-preconditioned LSMR on J itself, rhs = -r, no verification.
// DenseNormalCholesky::information_factor (src/normal.rs:341-392) and
// SelectedCovariance::compute_factor (src/covariance.rs:214-249) on a dense information
// matrix (column-major n x n, lower triangle significant; overwritten by its factor).
int orc_dense_covariance(const double* information, size_t n, const size_t* columns, size_t k, double tolerance,
                         double* out) {
  if (!is_finite(tolerance) || tolerance < 0.0 || tolerance >= 1.0) return ORC_INVALID_RANK_TOLERANCE;
  for (size_t a = 0; a < k; ++a)
    if (columns[a] >= n) return ORC_UNKNOWN_KEY;
  // row-major lower copy: row i of L is contiguous, so every Cholesky dot product streams
  std::vector<double> l(n * n, 0.0), diagonal(n);
  // factorize: a nonfinite lower entry is an evaluation error, mapped to InvalidInformation
  for (size_t c = 0; c < n; ++c) {
    diagonal[c] = information[c * n + c];
    for (size_t r = c; r < n; ++r) {
      const double v = information[c * n + r];
      if (!is_finite(v)) return ORC_INVALID_INFORMATION;
      l[r * n + c] = v;
    }
  }
  // column Cholesky (the cholesky_lower recurrence, same summation order), row-major storage
  for (size_t j = 0; j < n; ++j) {
    const double* lj = l.data() + j * n;
    double d = lj[j];
    for (size_t k = 0; k < j; ++k) d -= lj[k] * lj[k];
    if (!(d > 0.0) || !is_finite(d)) return ORC_SINGULAR_INFORMATION;
    d = std::sqrt(d);
    l[j * n + j] = d;
    for (size_t i = j + 1; i < n; ++i) {
      double* li = l.data() + i * n;
      double s = li[j];
      for (size_t k = 0; k < j; ++k) s -= li[k] * lj[k];
      li[j] = s / d;
    }
  }
  for (size_t i = 0; i < n; ++i) {
    const double pivot = l[i * n + i], d = diagonal[i];
    if (!is_finite(pivot) || !is_finite(d)) return ORC_INVALID_INFORMATION;
    const double ratio = pivot / std::sqrt(d);
    if (d <= 0.0 || pivot <= 0.0 || ratio * ratio <= tolerance) return ORC_SINGULAR_INFORMATION;
  }
  // L Y = E, covariance = Y^T Y
  std::vector<double> y(n * k, 0.0);
  for (size_t a = 0; a < k; ++a) {
    double* ya = y.data() + a * n;
    ya[columns[a]] = 1.0;
    for (size_t i = 0; i < n; ++i) {  // forward substitution with row i of L
      const double* li = l.data() + i * n;
      double s = ya[i];
      for (size_t t = 0; t < i; ++t) s -= li[t] * ya[t];
      ya[i] = s / li[i];
    }
  }
  for (size_t b = 0; b < k; ++b)
    for (size_t a = 0; a < k; ++a) {
      double s = 0.0;
      for (size_t i = 0; i < n; ++i) s += y[a * n + i] * y[b * n + i];
      if (!is_finite(s)) return ORC_INVALID_INFORMATION;
      out[b * k + a] = s;
    }
  return ORC_OK;
}

int orc_schur_covariance(const orc_schur* s, const size_t* offsets, const size_t* widths, size_t nblocks,
                         double tolerance, double* out) {
  if (!is_finite(tolerance) || tolerance < 0.0 || tolerance >= 1.0) return ORC_INVALID_RANK_TOLERANCE;
  const Jacobian& jac = s->model.jacobian;
  const size_t n = jac.cols;
  // selection: each (offset, width) must be a variable block of the model (the grid of the
  // Schur partition); SelectedCovariance::prepare (src/covariance.rs:180-207)
  std::vector<size_t> cols;
  for (size_t i = 0; i < nblocks; ++i) {
    const size_t o = offsets[i], w = widths[i];
    for (size_t j = 0; j < i; ++j)
      if (offsets[j] == o) return ORC_DUPLICATE_COVARIANCE_BLOCK;
    const bool ret = o < s->retained && o % s->rw == 0 && w == s->rw && o + w <= s->retained;
    const bool elim = o >= s->retained && (o - s->retained) % s->ew == 0 && w == s->ew && o + w <= n;
    if (!ret && !elim) return ORC_UNKNOWN_KEY;
    for (size_t c = o; c < o + w; ++c) cols.push_back(c);
  }
  // H = J^T J over the cached blocks, grouped by rows (src/linear_model.rs:122-167; only the
  // lower triangle is read)
  std::vector<double> h(n * n, 0.0);
  // blocks of one emission are contiguous (same row and row count)
  for (size_t a = 0; a < jac.blocks.size(); ++a) {
    const Block& A = jac.blocks[a];
    size_t lo = a, hi = a + 1;
    while (lo > 0 && jac.blocks[lo - 1].row == A.row && jac.blocks[lo - 1].rows == A.rows) --lo;
    while (hi < jac.blocks.size() && jac.blocks[hi].row == A.row && jac.blocks[hi].rows == A.rows) ++hi;
    for (size_t b = lo; b < hi; ++b) {
      const Block& B = jac.blocks[b];
      for (size_t ca = 0; ca < A.cols; ++ca)
        for (size_t cb = 0; cb < B.cols; ++cb) {
          const size_t r = A.col + ca, c = B.col + cb;
          if (r < c) continue;
          double v = 0.0;
          for (size_t i = 0; i < A.rows; ++i)
            v += jac.values[A.start + ca * A.rows + i] * jac.values[B.start + cb * B.rows + i];
          h[c * n + r] += v;
        }
    }
  }
  return orc_dense_covariance(h.data(), n, cols.data(), cols.size(), tolerance, out);
}

int orc_lsmr_solve(orc_schur* s, size_t max_iterations, double tol, double* delta, orc_lsmr_diagnostics* d) {
  *d = orc_lsmr_diagnostics{ORC_LSMR_NONE, 0, 0, kNaN, kNaN, kNaN, kNaN, kNaN, kNaN, kNaN, kNaN, kNaN, 0};
  if (max_iterations == 0 || !(tol >= 0.0 && tol < 1.0)) return ORC_INVALID_OPTIONS;
  const LinearizedModel& m = s->model;
  const size_t n = m.jacobian.cols, rows = m.jacobian.rows;
  std::fill(delta, delta + n, 0.0);
  if (rows == 0 || n == 0) {
    d->termination = ORC_LSMR_CONVERGED;
    d->finite_step = 1;
    return ORC_OK;
  }
  s->solves += 1;
  LsmrOperator op{s, 0, true, 0.0};
  LsmrResult r = lsmr(op, m.negative_residual.data(), delta, tol, max_iterations);
  s->iterations += r.iterations;
  d->iterations = r.iterations;
  d->estimated_absolute_residual = r.abs_residual;
  d->estimated_relative_residual = r.rel_residual;
  d->finite_step = norm_inf(delta, n, nullptr) == ORC_OK;
  const bool finite = d->finite_step && std::isfinite(r.abs_residual) && std::isfinite(r.rel_residual);
  d->termination = finite ? (r.converged ? ORC_LSMR_CONVERGED : ORC_LSMR_ITERATION_LIMIT) : ORC_LSMR_NON_FINITE;
  return d->termination == ORC_LSMR_CONVERGED ? ORC_OK : ORC_LINEAR_SOLVE_FAILED;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// BA problem and batch evaluator: examples/bal/model.rs:194-237, examples/bal.rs:176-201
// ---------------------------------------------------------------------------------------
struct orc_ba {
  size_t nc, np, no;
  std::vector<double> cams, pts, trial_cams, trial_pts;
  std::vector<int64_t> obs_cam, obs_pt;
  std::vector<double> obs_px;
  double huber = 0.0;  // 0 = plain least squares
  std::vector<int64_t> order, cam_off;  // rows in camera-stable order (Frame batches)
  std::vector<size_t> pt_off, pt_rows;   // each point's rows, ascending (its emission order)
  bool trial_active = false;
  orc_schur* backend = nullptr;  // used by the LM callbacks

  // Blockwise Huber (src/factors.rs:42-52): cost and sqrt(weight) for residual r.
  void robust(const double r[2], double* cost, double* scale) const {
    double s2 = r[0] * r[0] + r[1] * r[1];
    if (huber <= 0.0 || s2 <= huber * huber) {
      *cost = 0.5 * s2;
      *scale = 1.0;
    } else {
      double s = std::sqrt(s2);
      *cost = huber * (s - 0.5 * huber);
      *scale = std::sqrt(huber / s);
    }
  }

  // src/optimization.rs:391-430 + Frame::cost (model.rs:201-213): per batch sum in payload
  // order, checked, then added to the total in batch order.
  // Batch sums are independent; the total adds them in batch order.
  int cost(const std::vector<double>& c, const std::vector<double>& p, double* out) const {
    std::vector<double> batches(nc, 0.0);
    std::vector<int> status(nc, ORC_OK);
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t cam = 0; cam < nc; ++cam) {
      if (cam_off[cam] == cam_off[cam + 1]) continue;
      Projection proj(&c[9 * cam]);
      double batch = 0.0;
      for (int64_t k = cam_off[cam]; k < cam_off[cam + 1]; ++k) {
        int64_t o = order[k];
        double px[2];
        int st = proj.value(&p[3 * obs_pt[o]], px);
        if (st != ORC_OK) {
          status[cam] = st;
          break;
        }
        double r[2] = {px[0] - obs_px[2 * o], px[1] - obs_px[2 * o + 1]};
        double cst, sc;
        robust(r, &cst, &sc);
        batch += cst;
      }
      if (status[cam] == ORC_OK && (!is_finite(batch) || batch < 0.0)) status[cam] = ORC_INVALID_EVALUATION;
      batches[cam] = batch;
    }
    double total = 0.0;
    for (size_t cam = 0; cam < nc; ++cam) {
      if (cam_off[cam] == cam_off[cam + 1]) continue;
      if (status[cam] != ORC_OK) return status[cam];
      total += batches[cam];
    }
    if (!is_finite(total) || total < 0.0) return ORC_INVALID_EVALUATION;
    *out = total;
    return ORC_OK;
  }

  // One observation's emission: weighted residual rw and the 24 Jacobian values (camera
  // block 2x9 then point block 2x3, column-major), as Frame::linearize emits it.
  int emission(const Projection& proj, int64_t o, double rw[2], double vals[24]) const {
    const double* x = &pts[3 * obs_pt[o]];
    double px[2];
    int st = proj.value(x, px);
    if (st != ORC_OK) return st;
    double r[2] = {px[0] - obs_px[2 * o], px[1] - obs_px[2 * o + 1]};
    double jc[18], jp[6];
    proj.jacobians(x, jc, jp);
    double cst, sc;
    robust(r, &cst, &sc);
    for (int i = 0; i < 18; ++i) vals[i] = sc * jc[i];
    for (int i = 0; i < 6; ++i) vals[18 + i] = sc * jp[i];
    rw[0] = sc * r[0];
    rw[1] = sc * r[1];
    for (int i = 0; i < 24; ++i)
      if (!is_finite(vals[i])) return ORC_INVALID_EVALUATION;
    if (!is_finite(rw[0]) || !is_finite(rw[1])) return ORC_INVALID_EVALUATION;
    return ORC_OK;
  }

  // Frame::linearize for every batch in bulk: the same values, blocks, edges and gradient as
  // the per-emission path below (emission k of the camera-stable order owns rows 2k, 2k+1,
  // blocks 2k, 2k+1 and values 24k..24k+24), with the gradient of a camera summed over its
  // rows and that of a point over its rows in ascending order, which is the order in which
  // the per-emission path adds them.
  int linearize(orc_schur* s) const {
    if (s->retained != 9 * nc || s->rw != 9 || s->ew != 3 || s->model.jacobian.cols != 9 * nc + 3 * np)
      return linearize_sequential(s);
    orc_schur_clear(s);
    LinearizedModel& m = s->model;
    m.jacobian.rows = 2 * no;
    m.jacobian.blocks.resize(2 * no);
    m.jacobian.values.resize(24 * no);
    m.negative_residual.resize(2 * no);
    s->edges.resize(no);
    std::vector<int> status(nc, ORC_OK);
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t cam = 0; cam < nc; ++cam) {
      if (cam_off[cam] == cam_off[cam + 1]) continue;
      Projection proj(&cams[9 * cam]);
      for (int64_t k = cam_off[cam]; k < cam_off[cam + 1]; ++k) {
        int64_t o = order[k];
        double rw[2];
        int st = emission(proj, o, rw, &m.jacobian.values[24 * k]);
        if (st != ORC_OK) {
          status[cam] = st;
          break;
        }
        size_t row = 2 * (size_t)k, pcol = 3 * (size_t)obs_pt[o];
        m.negative_residual[row] = -rw[0];
        m.negative_residual[row + 1] = -rw[1];
        m.jacobian.blocks[row] = Block{row, 9 * cam, 2, 9, 24 * (size_t)k};
        m.jacobian.blocks[row + 1] = Block{row, 9 * nc + pcol, 2, 3, 24 * (size_t)k + 18};
        s->edges[k] = Edge{9 * cam, pcol, row, row + 1};
      }
    }
    for (size_t cam = 0; cam < nc; ++cam)
      if (status[cam] != ORC_OK) return status[cam];
    // gradient[col + j] += sum_i J_ij r_i per block (src/linear_model.rs:199-207)
    auto add = [&](size_t k, size_t first, size_t width, double* g) {
      double r[2] = {-m.negative_residual[2 * k], -m.negative_residual[2 * k + 1]};
      const double* v = &m.jacobian.values[24 * k + first];
      for (size_t j = 0; j < width; ++j) {
        double sum = 0;
        for (size_t i = 0; i < 2; ++i) sum += v[j * 2 + i] * r[i];
        g[j] += sum;
      }
    };
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t cam = 0; cam < nc; ++cam)
      for (int64_t k = cam_off[cam]; k < cam_off[cam + 1]; ++k) add(k, 0, 9, &m.gradient[9 * cam]);
#pragma omp parallel for schedule(dynamic, 256)
    for (size_t p = 0; p < np; ++p)
      for (size_t i = pt_off[p]; i < pt_off[p + 1]; ++i) add(pt_rows[i], 18, 3, &m.gradient[9 * nc + 3 * p]);
    return ORC_OK;
  }

  // Frame::linearize (model.rs:214-236) through the sink checks (optimization.rs:541-613),
  // one emission at a time.
  int linearize_sequential(orc_schur* s) const {
    orc_schur_clear(s);
    for (size_t cam = 0; cam < nc; ++cam) {
      if (cam_off[cam] == cam_off[cam + 1]) continue;
      Projection proj(&cams[9 * cam]);
      for (int64_t k = cam_off[cam]; k < cam_off[cam + 1]; ++k) {
        int64_t o = order[k];
        const double* x = &pts[3 * obs_pt[o]];
        double px[2];
        int st = proj.value(x, px);
        if (st != ORC_OK) return st;
        double r[2] = {px[0] - obs_px[2 * o], px[1] - obs_px[2 * o + 1]};
        double jc[18], jp[6];
        proj.jacobians(x, jc, jp);
        double cst, sc;
        robust(r, &cst, &sc);
        double vals[24];
        for (int i = 0; i < 18; ++i) vals[i] = sc * jc[i];
        for (int i = 0; i < 6; ++i) vals[18 + i] = sc * jp[i];
        double rw[2] = {sc * r[0], sc * r[1]};
        for (double v : vals)
          if (!is_finite(v)) return ORC_INVALID_EVALUATION;
        if (!is_finite(rw[0]) || !is_finite(rw[1])) return ORC_INVALID_EVALUATION;
        size_t cols[2] = {9 * cam, 9 * nc + 3 * (size_t)obs_pt[o]};
        size_t widths[2] = {9, 3};
        st = orc_schur_accumulate(s, rw, 2, 2, cols, widths, vals);
        if (st != ORC_OK) return st;
      }
    }
    return ORC_OK;
  }
};

extern "C" {

orc_ba* orc_ba_new(size_t nc, size_t np, size_t no, const double* cameras, const double* points,
                   const int64_t* obs_camera, const int64_t* obs_point, const double* obs_pixel,
                   double huber_delta) {
  orc_ba* b = new orc_ba();
  b->nc = nc;
  b->np = np;
  b->no = no;
  b->cams.assign(cameras, cameras + 9 * nc);
  b->pts.assign(points, points + 3 * np);
  b->obs_cam.assign(obs_camera, obs_camera + no);
  b->obs_pt.assign(obs_point, obs_point + no);
  b->obs_px.assign(obs_pixel, obs_pixel + 2 * no);
  b->huber = huber_delta;
  // Frame batch per camera, payloads in insertion order (examples/bal.rs:189-201).
  b->cam_off.assign(nc + 1, 0);
  for (size_t o = 0; o < no; ++o) b->cam_off[obs_camera[o] + 1]++;
  for (size_t c = 0; c < nc; ++c) b->cam_off[c + 1] += b->cam_off[c];
  b->order.assign(no, 0);
  std::vector<int64_t> fill(b->cam_off.begin(), b->cam_off.end() - 1);
  for (size_t o = 0; o < no; ++o) b->order[fill[obs_camera[o]]++] = (int64_t)o;
  std::vector<size_t> key(no);
  for (size_t k = 0; k < no; ++k) key[k] = (size_t)obs_point[b->order[k]];
  bucket(np, no, key.data(), b->pt_off, b->pt_rows);
  return b;
}
void orc_ba_free(orc_ba* b) { delete b; }
void orc_ba_layout(const orc_ba* b, int64_t* cam_order, int64_t* cam_offsets, int64_t* pt_perm,
                   int64_t* pt_offsets) {
  std::copy(b->order.begin(), b->order.end(), cam_order);
  std::copy(b->cam_off.begin(), b->cam_off.end(), cam_offsets);
  // Stable order by point of the camera-ordered rows: the order in which block_factors.rs
  // walks each point's emissions.
  std::vector<int64_t> off(b->np + 1, 0);
  for (size_t k = 0; k < b->no; ++k) off[b->obs_pt[b->order[k]] + 1]++;
  for (size_t p = 0; p < b->np; ++p) off[p + 1] += off[p];
  std::copy(off.begin(), off.end(), pt_offsets);
  std::vector<int64_t> fill(off.begin(), off.end() - 1);
  for (size_t k = 0; k < b->no; ++k) pt_perm[fill[b->obs_pt[b->order[k]]]++] = (int64_t)k;
}
int orc_ba_cost(const orc_ba* b, double* out) { return b->cost(b->cams, b->pts, out); }
void orc_ba_get_state(const orc_ba* b, double* cameras, double* points) {
  std::copy(b->cams.begin(), b->cams.end(), cameras);
  std::copy(b->pts.begin(), b->pts.end(), points);
}
void orc_ba_set_state(orc_ba* b, const double* cameras, const double* points) {
  b->cams.assign(cameras, cameras + 9 * b->nc);
  b->pts.assign(points, points + 3 * b->np);
}
int orc_ba_linearize(orc_ba* b, orc_schur* s) { return b->linearize(s); }
int orc_ba_linearize_sequential(orc_ba* b, orc_schur* s) { return b->linearize_sequential(s); }
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_get_threads(void) { return omp_get_max_threads(); }
void orc_ba_retract(orc_ba* b, const double* delta) {
  for (size_t i = 0; i < 9 * b->nc; ++i) b->cams[i] += delta[i];
  for (size_t i = 0; i < 3 * b->np; ++i) b->pts[i] += delta[9 * b->nc + i];
}

// ---------------------------------------------------------------------------------------
// LM: src/optimization/levenberg_marquardt.rs:330-472
// ---------------------------------------------------------------------------------------
static int validate_optimize(const orc_optimize_options* o) {  // src/optimization.rs:69-84
  if (o->max_iterations == 0) return ORC_INVALID_OPTIONS;
  for (double x : {o->gradient_tolerance, o->step_tolerance, o->cost_tolerance})
    if (!is_finite(x) || x < 0.0) return ORC_INVALID_OPTIONS;
  return ORC_OK;
}
static int validate_lm(const orc_lm_options* o) {  // levenberg_marquardt.rs:47-67
  if (o->max_trials == 0) return ORC_INVALID_OPTIONS;
  for (double x : {o->initial_damping, o->min_damping, o->max_damping, o->min_column_norm,
                   o->acceptance_threshold})
    if (!is_finite(x) || x <= 0.0) return ORC_INVALID_OPTIONS;
  if (o->min_damping > o->initial_damping || o->initial_damping > o->max_damping ||
      o->acceptance_threshold >= 1.0)
    return ORC_INVALID_OPTIONS;
  return ORC_OK;
}
static int checked_cost(int st, double value) {  // src/storage.rs:612-618
  if (st != ORC_OK) return st;
  if (!is_finite(value) || value < 0.0) return ORC_INVALID_EVALUATION;
  return ORC_OK;
}

static void init_outcome(orc_lm_outcome* out) {
  std::memset(out, 0, sizeof(*out));
  out->status = ORC_OK;
  out->termination = ORC_TERM_NONE;
  out->cost = kNaN;
  out->gradient_norm = kNaN;
  out->step_norm = kNaN;
  out->last_rejection = ORC_REJ_NONE;
  out->initial_cost = kNaN;
  out->final_cost = kNaN;
}

void orc_lm_optimize(const orc_problem* pb, const orc_optimize_options* options,
                     const orc_lm_options* lm, orc_lm_outcome* out) {
  init_outcome(out);
  auto finish = [&](int status) {
    out->status = status;
    if (pb->linear_work && out->attempts > 0) pb->linear_work(pb->ctx, &out->linear_solves, &out->linear_iterations);
  };
  int st = validate_optimize(options);
  if (st == ORC_OK) st = validate_lm(lm);
  if (st != ORC_OK) return finish(st);
  double lambda = lm->initial_damping;
  out->damping = lambda;
  st = pb->prepare(pb->ctx);
  if (st != ORC_OK) return finish(st);
  out->cost_evaluations += 1;
  double initial = 0;
  st = pb->cost(pb->ctx, 0, &initial);
  st = checked_cost(st, initial);
  if (st != ORC_OK) return finish(st);
  out->cost = initial;
  out->initial_cost = initial;
  double current = initial;
  bool stalled = false, tried_undamped = false;
  for (;;) {
    if (pb->dimension == 0) {
      out->final_cost = current;
      out->iterations = out->accepted_steps;
      out->termination = ORC_TERM_NO_VARIABLES;
      return finish(ORC_OK);
    }
    out->linearizations += 1;
    st = pb->linearize(pb->ctx);
    if (st != ORC_OK) return finish(st);
    double gradient = 0;
    st = pb->gradient_norm(pb->ctx, &gradient);
    if (st != ORC_OK) return finish(st);
    out->gradient_norm = gradient;
    if (gradient <= options->gradient_tolerance) {
      out->final_cost = current;
      out->iterations = out->accepted_steps;
      out->termination = ORC_TERM_GRADIENT;
      return finish(ORC_OK);
    }
    if (stalled && (tried_undamped || out->accepted_steps == options->max_iterations))
      return finish(ORC_NO_PROGRESS);
    if (out->accepted_steps == options->max_iterations) return finish(ORC_NO_CONVERGENCE);
    st = pb->prepare_damping(pb->ctx, lm->min_column_norm);
    if (st != ORC_OK) return finish(st);
    bool undamped_trial = stalled;
    tried_undamped |= undamped_trial;
    double growth = 2.0;
    bool accepted = false;
    for (size_t trial = 0; trial < lm->max_trials; ++trial) {
      out->attempts += 1;
      double damping = undamped_trial ? 0.0 : lambda;
      out->damping = damping;
      const double* delta = nullptr;
      double predicted = 0;
      int rejection;
      st = pb->solve_damped(pb->ctx, damping, &delta, &predicted);
      if (st == ORC_LINEAR_SOLVE_FAILED) {
        rejection = ORC_REJ_LINEAR_SOLVE;
      } else if (st != ORC_OK) {
        return finish(st);
      } else {
        double norm = 0;
        if (norm_inf(delta, pb->dimension, &norm) != ORC_OK) return finish(ORC_LINEAR_SOLVE_FAILED);
        out->step_norm = norm;
        if (!is_finite(predicted) || predicted <= 0.0) {
          rejection = ORC_REJ_PREDICTION;
        } else {
          st = pb->stage(pb->ctx, delta);
          double value = 0;
          if (st == ORC_OK) {
            out->cost_evaluations += 1;
            st = pb->cost(pb->ctx, 1, &value);
            st = checked_cost(st, value);
          }
          if (st == ORC_INVALID_EVALUATION) {
            pb->reject(pb->ctx);
            rejection = ORC_REJ_INVALID_EVALUATION;
          } else if (st != ORC_OK) {
            pb->reject(pb->ctx);
            return finish(st);
          } else {
            double actual = current - value;
            double ratio = std::min(actual / predicted, 2.0);
            if (actual > 0.0 && ratio >= lm->acceptance_threshold) {
              pb->accept(pb->ctx);
              out->accepted_steps += 1;
              out->cost = value;
              stalled = options->step_tolerance > 0.0 && norm <= options->step_tolerance &&
                        (options->cost_tolerance == 0.0 ||
                         actual / std::max(current, 1.0) <= options->cost_tolerance);
              current = value;
              // f64::powi(3) is repeated multiplication (x * x) * x
              double c = 2.0 * std::min(ratio, 1.0) - 1.0;
              double update = std::max(1.0 - c * c * c, 1.0 / 3.0);
              lambda = std::min(std::max(lambda * update, lm->min_damping), lm->max_damping);
              accepted = true;
              if (pb->on_accept) pb->on_accept(pb->ctx, out);
              break;
            }
            pb->reject(pb->ctx);
            rejection = ORC_REJ_POOR_AGREEMENT;
          }
        }
      }
      out->rejected_steps += 1;
      out->last_rejection = rejection;
      if (undamped_trial || lambda == lm->max_damping) break;
      lambda = std::min(lambda * growth, lm->max_damping);
      growth = std::min(growth * 2.0, std::max(lm->max_damping, 2.0));
    }
    if (!accepted) return finish(ORC_NO_PROGRESS);
  }
}

// src/optimization/gauss_newton.rs:101-176
void orc_gn_optimize(const orc_problem* pb, const orc_optimize_options* options, orc_lm_outcome* out) {
  init_outcome(out);
  auto finish = [&](int status) {
    out->status = status;
    if (pb->linear_work) pb->linear_work(pb->ctx, &out->linear_solves, &out->linear_iterations);
  };
  int st = validate_optimize(options);
  if (st != ORC_OK) return finish(st);
  st = pb->prepare(pb->ctx);
  if (st != ORC_OK) return finish(st);
  double initial = 0;
  st = checked_cost(pb->cost(pb->ctx, 0, &initial), initial);
  if (st != ORC_OK) return finish(st);
  out->initial_cost = initial;
  double current = initial;
  size_t iterations = 0;
  if (pb->dimension == 0) {
    out->final_cost = current;
    out->termination = ORC_TERM_NO_VARIABLES;
    return finish(ORC_OK);
  }
  for (;;) {
    out->linearizations += 1;
    st = pb->linearize(pb->ctx);
    if (st != ORC_OK) return finish(st);
    double gradient = 0;
    st = pb->gradient_norm(pb->ctx, &gradient);
    if (st != ORC_OK) return finish(st);
    out->gradient_norm = gradient;
    if (gradient <= options->gradient_tolerance) {
      out->final_cost = current;
      out->iterations = iterations;
      out->termination = ORC_TERM_GRADIENT;
      return finish(ORC_OK);
    }
    if (iterations == options->max_iterations) return finish(ORC_NO_CONVERGENCE);
    const double* delta = nullptr;
    st = pb->solve(pb->ctx, &delta);
    if (st != ORC_OK) return finish(st);
    double step_norm = 0;
    if (norm_inf(delta, pb->dimension, &step_norm) != ORC_OK) return finish(ORC_LINEAR_SOLVE_FAILED);
    out->step_norm = step_norm;
    if (step_norm <= options->step_tolerance) {
      out->final_cost = current;
      out->iterations = iterations;
      out->termination = ORC_TERM_STEP;
      return finish(ORC_OK);
    }
    st = pb->stage(pb->ctx, delta);
    double trial = 0;
    if (st == ORC_OK) st = checked_cost(pb->cost(pb->ctx, 1, &trial), trial);
    if (st != ORC_OK) {
      pb->reject(pb->ctx);
      return finish(st);
    }
    pb->accept(pb->ctx);
    iterations += 1;
    out->accepted_steps = iterations;
    double decrease = current - trial;
    bool converged = trial == 0.0 ||
                     (decrease >= 0.0 && decrease / std::max(current, 1.0) <= options->cost_tolerance);
    current = trial;
    out->cost = current;
    if (converged) {
      out->final_cost = current;
      out->iterations = iterations;
      out->termination = ORC_TERM_COST;
      return finish(ORC_OK);
    }
  }
}

// BA problem adapter for the generic loops.
namespace {
struct BaCtx {
  orc_ba* ba;
  orc_schur* s;
  std::vector<double> delta;
  double* trace;
  size_t trace_cap, trace_len;
};
int ba_prepare(void* c) {
  BaCtx* x = (BaCtx*)c;
  x->ba->trial_active = false;
  return orc_schur_prepare(x->s, 9 * x->ba->nc + 3 * x->ba->np);
}
int ba_cost(void* c, int trial, double* out) {
  BaCtx* x = (BaCtx*)c;
  if (trial) return x->ba->cost(x->ba->trial_cams, x->ba->trial_pts, out);
  return x->ba->cost(x->ba->cams, x->ba->pts, out);
}
int ba_linearize(void* c) { BaCtx* x = (BaCtx*)c; return x->ba->linearize(x->s); }
int ba_gradient_norm(void* c, double* out) { return orc_schur_gradient_norm(((BaCtx*)c)->s, out); }
int ba_prepare_damping(void* c, double floor) { return orc_schur_prepare_damping(((BaCtx*)c)->s, floor); }
int ba_solve_damped(void* c, double lambda, const double** delta, double* pred) {
  BaCtx* x = (BaCtx*)c;
  x->delta.resize(orc_schur_dimension(x->s));
  int st = orc_schur_solve_damped(x->s, lambda, x->delta.data(), pred);
  *delta = x->delta.data();
  return st;
}
int ba_solve(void* c, const double** delta) {
  BaCtx* x = (BaCtx*)c;
  x->delta.resize(orc_schur_dimension(x->s));
  int st = orc_schur_solve(x->s, x->delta.data());
  *delta = x->delta.data();
  return st;
}
int ba_stage(void* c, const double* delta) {
  orc_ba* b = ((BaCtx*)c)->ba;
  b->trial_cams = b->cams;
  b->trial_pts = b->pts;
  size_t nc9 = 9 * b->nc, np3 = 3 * b->np;
  for (size_t i = 0; i < nc9; ++i) b->trial_cams[i] += delta[i];
#pragma omp parallel for schedule(static)
  for (size_t i = 0; i < np3; ++i) b->trial_pts[i] += delta[nc9 + i];
  b->trial_active = true;
  return ORC_OK;
}
void ba_accept(void* c) {
  orc_ba* b = ((BaCtx*)c)->ba;
  if (b->trial_active) {
    std::swap(b->cams, b->trial_cams);
    std::swap(b->pts, b->trial_pts);
    b->trial_active = false;
  }
}
void ba_reject(void* c) { ((BaCtx*)c)->ba->trial_active = false; }
void ba_work(void* c, size_t* solves, size_t* iterations) {
  orc_schur_statistics(((BaCtx*)c)->s, solves, iterations);
}
// One record per accepted step: cost, attempts so far, damping of the accepted trial.
void ba_on_accept(void* c, const orc_lm_outcome* st) {
  BaCtx* x = (BaCtx*)c;
  if (x->trace && x->trace_len < x->trace_cap) {
    double* r = x->trace + 3 * x->trace_len++;
    r[0] = st->cost;
    r[1] = (double)st->attempts;
    r[2] = st->damping;
  }
}
}  // namespace

void orc_ba_lm(orc_ba* ba, size_t max_cg, double tol, const orc_optimize_options* options,
               const orc_lm_options* lm, orc_lm_outcome* out, double* trace, size_t cap, size_t* len) {
  orc_schur* s = orc_schur_new(9 * ba->nc, 9, 3);
  orc_schur_set_options(s, max_cg, tol);
  BaCtx ctx{ba, s, {}, trace, cap, 0};
  orc_problem pb{&ctx, 9 * ba->nc + 3 * ba->np, ba_prepare, ba_cost, ba_linearize, ba_gradient_norm,
                 ba_prepare_damping, ba_solve_damped, ba_solve, ba_stage, ba_accept, ba_reject,
                 ba_work, ba_on_accept};
  orc_lm_optimize(&pb, options, lm, out);
  if (len) *len = ctx.trace_len;
  orc_schur_free(s);
}
void orc_ba_gn(orc_ba* ba, size_t max_cg, double tol, const orc_optimize_options* options,
               orc_lm_outcome* out) {
  orc_schur* s = orc_schur_new(9 * ba->nc, 9, 3);
  orc_schur_set_options(s, max_cg, tol);
  BaCtx ctx{ba, s, {}, nullptr, 0, 0};
  orc_problem pb{&ctx, 9 * ba->nc + 3 * ba->np, ba_prepare, ba_cost, ba_linearize, ba_gradient_norm,
                 ba_prepare_damping, ba_solve_damped, ba_solve, ba_stage, ba_accept, ba_reject,
                 ba_work, nullptr};
  orc_gn_optimize(&pb, options, out);
  orc_schur_free(s);
}

}  // extern "C"
