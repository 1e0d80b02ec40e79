// C++ drop-ins for NetworKit's CSRGeneralMatrix<double> product and AlgebraicPageRank on the
// MI355X library (nkpr_csr_spmv, nkpr_algebraic_pagerank).
//
// Interface kept from the reference (kit-parco/networkit v9.0):
//   CSRGeneralMatrix   include/networkit/algebraic/CSRGeneralMatrix.hpp: the triplet
//                      constructors (:138-182), adjacencyMatrix (:889-900), numberOfRows /
//                      numberOfColumns / nnz, forNonZeroElementsInRowOrder, operator*(Vector)
//                      (:563-578); CSRMatrix = CSRGeneralMatrix<double> (CSRMatrix.hpp:10)
//   Vector             include/networkit/algebraic/Vector.hpp: the parts a product needs
//                      (dimension, element access, transposed flag)
//   AlgebraicPageRank  include/networkit/algebraic/algorithms/AlgebraicPageRank.hpp:25-150
//                      (constructor (G, damp = 0.85, tol = 1e-8), run, scores, ranking, score,
//                      maximum)
// Only double values are supported.  The matrix keeps a device copy of its arrays, made on the
// first product and kept until the matrix is destroyed (the class has no mutating members).
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <utility>
#include <vector>

#include "../../include/nkpr.h"
#include "Graph.hpp"
#include "PageRank.hpp"

namespace NetworKit {

struct Triplet {
    index row;
    index column;
    double value;
};

class Vector {
public:
    Vector() = default;
    explicit Vector(count dimension, double initialValue = 0.0, bool transpose = false)
        : values(dimension, initialValue), transposed(transpose) {}
    Vector(std::vector<double> v, bool transpose = false)
        : values(std::move(v)), transposed(transpose) {}
    count getDimension() const noexcept { return values.size(); }
    bool isTransposed() const noexcept { return transposed; }
    double &operator[](index i) { return values[i]; }
    double operator[](index i) const { return values[i]; }
    double &at(index i) { return values.at(i); }
    const std::vector<double> &data() const noexcept { return values; }
    std::vector<double> &data() noexcept { return values; }

private:
    std::vector<double> values;
    bool transposed = false;
};

// Device copy of a CSR matrix's three arrays.
class DeviceCsrArrays {
public:
    DeviceCsrArrays() = default;
    ~DeviceCsrArrays();
    DeviceCsrArrays(const DeviceCsrArrays &) = delete;
    DeviceCsrArrays &operator=(const DeviceCsrArrays &) = delete;
    void upload(const std::vector<index> &rowIdx, const std::vector<index> &columnIdx,
                const std::vector<double> &values);
    nkpr_csr view(count rows, count cols, count nnz) const;
    bool ready() const noexcept { return ok; }

private:
    void *row = nullptr, *col = nullptr, *val = nullptr;
    bool ok = false;
};

template <typename ValueType> class CSRGeneralMatrix;

template <> class CSRGeneralMatrix<double> {
public:
    CSRGeneralMatrix() = default;
    CSRGeneralMatrix(count dimension, const std::vector<Triplet> &triplets, double zero = 0,
                     bool isSorted = false)
        : CSRGeneralMatrix(dimension, dimension, triplets, zero, isSorted) {}
    CSRGeneralMatrix(count nRows, count nCols, const std::vector<Triplet> &triplets,
                     double zero = 0, bool isSorted = false);
    CSRGeneralMatrix(count nRows, count nCols, std::vector<index> rowIdx,
                     std::vector<index> columnIdx, std::vector<double> nonZeros, double zero = 0,
                     bool isSorted = false);
    CSRGeneralMatrix(CSRGeneralMatrix &&) = default;
    CSRGeneralMatrix &operator=(CSRGeneralMatrix &&) = default;

    static CSRGeneralMatrix adjacencyMatrix(const Graph &graph, double zero = 0);

    count numberOfRows() const noexcept { return nRows; }
    count numberOfColumns() const noexcept { return nCols; }
    count nnz() const noexcept { return nonZeros.size(); }
    double getZero() const noexcept { return zero; }
    bool sorted() const noexcept { return isSorted; }

    template <typename L> void forNonZeroElementsInRowOrder(L handle) const {
        for (index i = 0; i < nRows; ++i)
            for (index k = rowIdx[i]; k < rowIdx[i + 1]; ++k) handle(i, columnIdx[k], nonZeros[k]);
    }
    template <typename L> void forNonZeroElementsInRow(index i, L handle) const {
        for (index k = rowIdx[i]; k < rowIdx[i + 1]; ++k) handle(columnIdx[k], nonZeros[k]);
    }

    // y = A x on the device (CSRGeneralMatrix.hpp:563-578).
    Vector operator*(const Vector &vector) const;

    // The device view of the matrix (uploaded on first use).
    nkpr_csr deviceView() const;

    const std::vector<index> &rowIndices() const noexcept { return rowIdx; }
    const std::vector<index> &columnIndices() const noexcept { return columnIdx; }
    const std::vector<double> &values() const noexcept { return nonZeros; }

private:
    std::vector<index> rowIdx{0};
    std::vector<index> columnIdx;
    std::vector<double> nonZeros;
    count nRows = 0, nCols = 0;
    bool isSorted = false;
    double zero = 0;
    mutable std::unique_ptr<DeviceCsrArrays> dev;
};

using CSRMatrix = CSRGeneralMatrix<double>;

// The row sums of a matrix (GraphBLAS::rowReduce, GraphBLAS.hpp:278-291).
Vector rowReduce(const CSRMatrix &matrix);

template <class Matrix> class AlgebraicPageRank final : public Centrality {
public:
    AlgebraicPageRank(const Graph &graph, double damp = 0.85, double tol = 1e-8);
    void run() override;
    const std::vector<double> &scores() const override;
    std::vector<std::pair<node, double>> ranking() override;
    double score(node v) override;
    double maximum() override { return 1.0; }
    count numberOfIterations() const {
        assureFinished();
        return iterations;
    }
    const Matrix &matrix() const noexcept { return M; }

private:
    Matrix M;
    const double damp;
    const double tol;
    count iterations = 0;
    // the device operator built from M on the first run and kept for the next ones
    std::shared_ptr<nkpr_algebraic> device;
};

template <> AlgebraicPageRank<CSRMatrix>::AlgebraicPageRank(const Graph &graph, double damp,
                                                            double tol);
template <> void AlgebraicPageRank<CSRMatrix>::run();
template <> const std::vector<double> &AlgebraicPageRank<CSRMatrix>::scores() const;
template <> std::vector<std::pair<node, double>> AlgebraicPageRank<CSRMatrix>::ranking();
template <> double AlgebraicPageRank<CSRMatrix>::score(node v);

}  // namespace NetworKit
