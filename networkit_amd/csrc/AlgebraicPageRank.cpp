// CSRGeneralMatrix<double> and AlgebraicPageRank<CSRMatrix> on the MI355X library.
//
// The matrix is built on the host exactly as the reference builds it (the triplet constructor
// is a counting sort that keeps the triplets' order inside a row, CSRGeneralMatrix.hpp:151-182);
// products run on the device through nkpr_csr_spmv, the PageRank loop through a handle built
// from M once (nkpr_algebraic_create / nkpr_algebraic_run).
#include "AlgebraicPageRank.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <string>

namespace NetworKit {

namespace {
void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + " failed: " + hipGetErrorString(e));
}
void nkpr_check(int rc) {
    if (rc != NKPR_OK) throw std::runtime_error(nkpr_last_error());
}
}  // namespace

DeviceCsrArrays::~DeviceCsrArrays() {
    if (row) (void)hipFree(row);
    if (col) (void)hipFree(col);
    if (val) (void)hipFree(val);
}

void DeviceCsrArrays::upload(const std::vector<index> &rowIdx, const std::vector<index> &columnIdx,
                             const std::vector<double> &values) {
    const size_t rb = rowIdx.size() * sizeof(index);
    const size_t cb = std::max<size_t>(columnIdx.size(), 1) * sizeof(index);
    const size_t vb = std::max<size_t>(values.size(), 1) * sizeof(double);
    hip_check(hipMalloc(&row, rb), "hipMalloc");
    hip_check(hipMalloc(&col, cb), "hipMalloc");
    hip_check(hipMalloc(&val, vb), "hipMalloc");
    hip_check(hipMemcpy(row, rowIdx.data(), rb, hipMemcpyHostToDevice), "hipMemcpy");
    if (!columnIdx.empty()) {
        hip_check(hipMemcpy(col, columnIdx.data(), columnIdx.size() * sizeof(index),
                            hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(val, values.data(), values.size() * sizeof(double),
                            hipMemcpyHostToDevice), "hipMemcpy");
    }
    ok = true;
}

nkpr_csr DeviceCsrArrays::view(count rows, count cols, count nnz) const {
    nkpr_csr v;
    v.rows = rows;
    v.cols = cols;
    v.nnz = nnz;
    v.row_idx = static_cast<const uint64_t *>(row);
    v.col_idx = static_cast<const uint64_t *>(col);
    v.values = static_cast<const double *>(val);
    return v;
}

CSRGeneralMatrix<double>::CSRGeneralMatrix(count nRows_, count nCols_,
                                           const std::vector<Triplet> &triplets, double zero_,
                                           bool isSorted_)
    : rowIdx(nRows_ + 1, 0), columnIdx(triplets.size()), nonZeros(triplets.size()), nRows(nRows_),
      nCols(nCols_), isSorted(isSorted_), zero(zero_) {
    // counting sort by row; a row keeps its triplets in their given order
    for (const Triplet &t : triplets) {
        if (t.row >= nRows || t.column >= nCols)
            throw std::out_of_range("triplet outside the matrix");
        ++rowIdx[t.row + 1];
    }
    for (index i = 0; i < nRows; ++i) rowIdx[i + 1] += rowIdx[i];
    std::vector<index> next(rowIdx.begin(), rowIdx.end() - 1);
    for (const Triplet &t : triplets) {
        const index d = next[t.row]++;
        columnIdx[d] = t.column;
        nonZeros[d] = t.value;
    }
}

CSRGeneralMatrix<double>::CSRGeneralMatrix(count nRows_, count nCols_, std::vector<index> rowIdx_,
                                           std::vector<index> columnIdx_,
                                           std::vector<double> nonZeros_, double zero_,
                                           bool isSorted_)
    : rowIdx(std::move(rowIdx_)), columnIdx(std::move(columnIdx_)),
      nonZeros(std::move(nonZeros_)), nRows(nRows_), nCols(nCols_), isSorted(isSorted_),
      zero(zero_) {
    if (rowIdx.size() != nRows + 1 || columnIdx.size() != nonZeros.size() ||
        rowIdx.back() != columnIdx.size())
        throw std::invalid_argument("inconsistent CSR arrays");
}

CSRGeneralMatrix<double> CSRGeneralMatrix<double>::adjacencyMatrix(const Graph &graph,
                                                                   double zero) {
    // CSRGeneralMatrix.hpp:889-900: one triplet per forEdges edge, its mirror when undirected
    std::vector<Triplet> triplets;
    triplets.reserve(graph.isDirected() ? graph.numberOfEdges() : 2 * graph.numberOfEdges());
    const bool directed = graph.isDirected();
    graph.forEdges([&](node i, node j, double w) {
        triplets.push_back({i, j, w});
        if (!directed && i != j) triplets.push_back({j, i, w});
    });
    return CSRGeneralMatrix(graph.upperNodeIdBound(), triplets, zero);
}

nkpr_csr CSRGeneralMatrix<double>::deviceView() const {
    if (!dev) {
        auto d = std::make_unique<DeviceCsrArrays>();
        d->upload(rowIdx, columnIdx, nonZeros);
        dev = std::move(d);
    }
    return dev->view(nRows, nCols, nonZeros.size());
}

Vector CSRGeneralMatrix<double>::operator*(const Vector &vector) const {
    if (vector.isTransposed()) throw std::invalid_argument("the vector must not be transposed");
    if (vector.getDimension() != nCols) throw std::invalid_argument("dimension mismatch");
    Vector result(nRows, zero);
    if (nRows == 0) return result;
    const nkpr_csr a = deviceView();
    void *x = nullptr, *y = nullptr;
    struct Free {
        void *&p;
        ~Free() {
            if (p) (void)hipFree(p);
        }
    } fx{x}, fy{y};
    hip_check(hipMalloc(&x, std::max<size_t>(nCols, 1) * sizeof(double)), "hipMalloc");
    hip_check(hipMalloc(&y, nRows * sizeof(double)), "hipMalloc");
    if (nCols)
        hip_check(hipMemcpy(x, vector.data().data(), nCols * sizeof(double),
                            hipMemcpyHostToDevice), "hipMemcpy");
    nkpr_check(nkpr_csr_spmv(&a, static_cast<const double *>(x), static_cast<double *>(y),
                             nullptr));
    hip_check(hipMemcpy(result.data().data(), y, nRows * sizeof(double), hipMemcpyDeviceToHost),
              "hipMemcpy");
    return result;
}

Vector rowReduce(const CSRMatrix &matrix) {
    Vector r(matrix.numberOfRows(), 0.0);
    for (index i = 0; i < matrix.numberOfRows(); ++i)
        matrix.forNonZeroElementsInRow(i, [&](index, double v) { r[i] += v; });
    return r;
}

template <>
AlgebraicPageRank<CSRMatrix>::AlgebraicPageRank(const Graph &graph, double damp_, double tol_)
    : Centrality(graph), damp(damp_), tol(tol_) {
    // AlgebraicPageRank.hpp:35-54: M(j, i) = damp * A(i, j) / rowsum_i(A), taken row by row
    const CSRMatrix A = CSRMatrix::adjacencyMatrix(graph);
    Vector invOutDeg = rowReduce(A);
    for (index i = 0; i < invOutDeg.getDimension(); ++i) invOutDeg[i] = 1.0 / invOutDeg[i];
    std::vector<Triplet> mTriplets;
    mTriplets.reserve(A.nnz());
    A.forNonZeroElementsInRowOrder([&](index i, index j, double value) {
        mTriplets.push_back({j, i, damp * value * invOutDeg[i]});
    });
    M = CSRMatrix(A.numberOfRows(), mTriplets);
}

template <> void AlgebraicPageRank<CSRMatrix>::run() {
    // AlgebraicPageRank.hpp:94-119 on the device: the first run builds a PageRank handle from M
    // (nkpr_algebraic_create) and every run iterates on it (nkpr_algebraic_run)
    const count n = M.numberOfRows();
    if (n == 0) throw std::runtime_error("AlgebraicPageRank on an empty graph");
    if (!device) {
        const nkpr_csr m = M.deviceView();
        nkpr_algebraic *h = nullptr;
        nkpr_check(nkpr_algebraic_create(&m, &h));
        device.reset(h, [](nkpr_algebraic *p) { nkpr_algebraic_destroy(p); });
    }
    scoreData.assign(n, 0.0);
    uint64_t it = 0;
    nkpr_check(nkpr_algebraic_run(device.get(), damp, tol, UINT64_MAX, scoreData.data(), 0, &it,
                                  nullptr));
    iterations = it;
    hasRun = true;
}

template <> const std::vector<double> &AlgebraicPageRank<CSRMatrix>::scores() const {
    assureFinished();
    return scoreData;
}

template <> std::vector<std::pair<node, double>> AlgebraicPageRank<CSRMatrix>::ranking() {
    // AlgebraicPageRank.hpp:126-135: every slot, score descending (the reference's sort is
    // not stable; equal scores keep id order here, one of the orders it may produce)
    assureFinished();
    std::vector<std::pair<node, double>> r;
    r.reserve(scoreData.size());
    for (index i = 0; i < scoreData.size(); ++i) r.emplace_back(i, scoreData[i]);
    std::stable_sort(r.begin(), r.end(), [](const std::pair<node, double> &x,
                                            const std::pair<node, double> &y) {
        // a NaN score ranks after every number, so that the order stays a strict weak order
        const bool nx = std::isnan(x.second), ny = std::isnan(y.second);
        if (nx || ny) return !nx && ny;
        return x.second > y.second;
    });
    return r;
}

template <> double AlgebraicPageRank<CSRMatrix>::score(node v) {
    assureFinished();
    return scoreData.at(v);
}

}  // namespace NetworKit
