// multiple_sheet.h -- UVLM::MultipleSheet<T>, a (num, rows, cols) lattice.
//
// The reference takes this container from its multiple_sheet submodule
// (absent from the reference tree).  The contract below is the one the
// reference's call sites rely on: contiguous storage, sheet-major and
// row-major, K = j + cols * (i + rows * n); list_index() yields (K, n, i, j)
// in K order; prepend_row() inserts one row per sheet as row 0 and
// erase_last_row() drops the last row of every sheet.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <tuple>
#include <vector>

namespace UVLM {

template <class T>
class MultipleSheet {
 public:
  typedef std::tuple<std::size_t, std::size_t, std::size_t, std::size_t> index_t;
  typedef typename std::vector<T>::iterator iterator;
  typedef typename std::vector<T>::const_iterator const_iterator;

  MultipleSheet() : num_(0), rows_(0), cols_(0) {}
  MultipleSheet(std::size_t num, std::size_t rows, std::size_t cols)
      : num_(num), rows_(rows), cols_(cols), data_(num * rows * cols) {}

  void resize(std::size_t num, std::size_t rows, std::size_t cols) {
    num_ = num;
    rows_ = rows;
    cols_ = cols;
    data_.resize(num * rows * cols);
  }
  template <class U>
  void resize(const MultipleSheet<U>& other) {
    resize(other.num(), other.rows(), other.cols());
  }

  std::size_t num() const { return num_; }
  std::size_t rows() const { return rows_; }
  std::size_t cols() const { return cols_; }
  std::size_t size() const { return data_.size(); }

  std::size_t index(std::size_t n, std::size_t i, std::size_t j) const {
    return j + cols_ * (i + rows_ * n);
  }
  T& at(std::size_t n, std::size_t i, std::size_t j) { return data_[index(n, i, j)]; }
  const T& at(std::size_t n, std::size_t i, std::size_t j) const {
    return data_[index(n, i, j)];
  }
  T& operator[](std::size_t K) { return data_[K]; }
  const T& operator[](std::size_t K) const { return data_[K]; }

  iterator begin() { return data_.begin(); }
  iterator end() { return data_.end(); }
  const_iterator begin() const { return data_.begin(); }
  const_iterator end() const { return data_.end(); }
  iterator iterator_at(std::size_t n, std::size_t i, std::size_t j) {
    return data_.begin() + index(n, i, j);
  }
  const_iterator iterator_at(std::size_t n, std::size_t i, std::size_t j) const {
    return data_.begin() + index(n, i, j);
  }
  iterator sheet_begin(std::size_t n) { return iterator_at(n, 0, 0); }
  iterator sheet_end(std::size_t n) { return iterator_at(n + 1, 0, 0); }
  const_iterator sheet_begin(std::size_t n) const { return iterator_at(n, 0, 0); }
  const_iterator sheet_end(std::size_t n) const { return iterator_at(n + 1, 0, 0); }

  T* data() { return data_.data(); }
  const T* data() const { return data_.data(); }

  std::vector<index_t> list_index() const {
    std::vector<index_t> res;
    res.reserve(size());
    std::size_t K = 0;
    for (std::size_t n = 0; n < num_; ++n)
      for (std::size_t i = 0; i < rows_; ++i)
        for (std::size_t j = 0; j < cols_; ++j) res.emplace_back(K++, n, i, j);
    return res;
  }
  std::vector<index_t> list_index(std::size_t n) const {
    std::vector<index_t> res;
    for (std::size_t i = 0; i < rows_; ++i)
      for (std::size_t j = 0; j < cols_; ++j) res.emplace_back(index(n, i, j), n, i, j);
    return res;
  }

  // Insert [first, last) (num * cols elements, one row per sheet in sheet
  // order) as row 0 of every sheet.
  template <class InputIterator>
  void prepend_row(InputIterator first, InputIterator last) {
    std::vector<T> row(first, last);
    if (row.size() != num_ * cols_)
      throw std::invalid_argument("prepend_row: expected num*cols elements");
    std::vector<T> nd;
    nd.reserve(num_ * (rows_ + 1) * cols_);
    for (std::size_t n = 0; n < num_; ++n) {
      nd.insert(nd.end(), row.begin() + n * cols_, row.begin() + (n + 1) * cols_);
      nd.insert(nd.end(), data_.begin() + n * rows_ * cols_,
                data_.begin() + (n + 1) * rows_ * cols_);
    }
    data_.swap(nd);
    rows_ += 1;
  }

  void erase_last_row() {
    if (rows_ == 0) return;
    std::vector<T> nd;
    nd.reserve(num_ * (rows_ - 1) * cols_);
    for (std::size_t n = 0; n < num_; ++n)
      nd.insert(nd.end(), data_.begin() + n * rows_ * cols_,
                data_.begin() + (n * rows_ + rows_ - 1) * cols_);
    data_.swap(nd);
    rows_ -= 1;
  }

 private:
  std::size_t num_, rows_, cols_;
  std::vector<T> data_;
};

}  // namespace UVLM
