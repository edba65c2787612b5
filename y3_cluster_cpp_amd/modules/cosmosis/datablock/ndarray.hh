// ndarray.hh -- cosmosis::ndarray<T>, a row-major n-dimensional array that
// owns a copy of its data (source-compatible subset of CosmoSIS's class:
// construction from a std::vector plus extents, operator() indexing,
// extents(), ndims(), size(), iteration).
#ifndef Y3G_SHIM_NDARRAY_HH
#define Y3G_SHIM_NDARRAY_HH

#include <cstddef>
#include <initializer_list>
#include <stdexcept>
#include <vector>

namespace cosmosis {

  template <typename T>
  class ndarray {
  public:
    ndarray() = default;

    ndarray(std::vector<T> const& data, std::vector<std::size_t> const& extents)
      : data_(data), extents_(extents)
    {
      if (count_(extents_) != data_.size())
        throw std::invalid_argument("ndarray: extents do not match data size");
    }

    ndarray(T const* data, int ndims, int const* extents)
    {
      if (ndims <= 0) throw std::invalid_argument("ndarray: ndims <= 0");
      for (int i = 0; i != ndims; ++i) {
        if (extents[i] < 0) throw std::invalid_argument("ndarray: negative extent");
        extents_.push_back(static_cast<std::size_t>(extents[i]));
      }
      data_.assign(data, data + count_(extents_));
    }

    std::size_t ndims() const { return extents_.size(); }
    std::vector<std::size_t> const& extents() const { return extents_; }
    std::size_t size() const { return data_.size(); }

    template <typename... I>
    T&
    operator()(I... idx)
    {
      return data_[offset_({static_cast<std::size_t>(idx)...})];
    }
    template <typename... I>
    T const&
    operator()(I... idx) const
    {
      return data_[offset_({static_cast<std::size_t>(idx)...})];
    }

    T* data() { return data_.data(); }
    T const* data() const { return data_.data(); }
    typename std::vector<T>::iterator begin() { return data_.begin(); }
    typename std::vector<T>::iterator end() { return data_.end(); }
    typename std::vector<T>::const_iterator begin() const { return data_.begin(); }
    typename std::vector<T>::const_iterator end() const { return data_.end(); }

    bool
    operator==(ndarray const& o) const
    {
      return extents_ == o.extents_ && data_ == o.data_;
    }

  private:
    static std::size_t
    count_(std::vector<std::size_t> const& e)
    {
      std::size_t n = 1;
      for (auto x : e) n *= x;
      return e.empty() ? 0 : n;
    }

    std::size_t
    offset_(std::initializer_list<std::size_t> idx) const
    {
      if (idx.size() != extents_.size())
        throw std::out_of_range("ndarray: wrong number of indices");
      std::size_t off = 0, k = 0;
      for (auto i : idx) {
        if (i >= extents_[k]) throw std::out_of_range("ndarray: index out of range");
        off = off * extents_[k] + i;
        ++k;
      }
      return off;
    }

    std::vector<T> data_;
    std::vector<std::size_t> extents_;
  };

} // namespace cosmosis

#endif
