// datablock.hh -- cosmosis::DataBlock, a source-compatible subset of the
// CosmoSIS parameter/product store used by the integration modules.
//
// A DataBlock maps (section, name) -- both case-insensitive -- to a typed
// value: int, double, bool, std::string, or an int/double/string array.
// The member templates used by the modules and by the reference's module
// code are provided with CosmoSIS semantics:
//   get_val<T>(section, name, T& out)          status, no type coercion
//   get_val<T>(section, name, T const& def, T& out)   default if missing
//   put_val<T>(section, name, T const& v)       fails if the name exists
//   replace_val<T>(section, name, T const& v)   fails if it does not
//   view<T>(section, name)                      throws BadDataBlockAccess
//   has_section / has_val
// One-dimensional arrays are readable both as std::vector<T> and as
// cosmosis::ndarray<T>.  The C API over the same object is c_datablock.h.
#ifndef Y3G_SHIM_DATABLOCK_HH
#define Y3G_SHIM_DATABLOCK_HH

#include "cosmosis/datablock/datablock_status.h"
#include "cosmosis/datablock/ndarray.hh"

#include <map>
#include <ostream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <variant>
#include <vector>

namespace cosmosis {

  class Exception : public std::runtime_error {
  public:
    using std::runtime_error::runtime_error;
  };

  class DataBlock {
  public:
    struct BadDataBlockAccess : Exception {
      BadDataBlockAccess(std::string const& section, std::string const& name,
                         DATABLOCK_STATUS s)
        : Exception("DataBlock: bad access to " + section + "/" + name +
                    " (status " + std::to_string(int(s)) + ")"),
          status(s)
      {}
      DATABLOCK_STATUS status;
    };

    using value_t = std::variant<int, double, bool, std::string, ndarray<int>,
                                 ndarray<double>, std::vector<std::string>>;

    template <typename T>
    DATABLOCK_STATUS get_val(std::string const& section, std::string const& name,
                             T& val) const;

    template <typename T>
    DATABLOCK_STATUS
    get_val(std::string const& section, std::string const& name, T const& def,
            T& val) const
    {
      if (!has_val(section, name)) {
        val = def;
        return DBS_USED_DEFAULT;
      }
      return get_val(section, name, val);
    }

    template <typename T>
    DATABLOCK_STATUS
    put_val(std::string const& section, std::string const& name, T const& val)
    {
      if (has_val(section, name)) return DBS_NAME_ALREADY_EXISTS;
      sections_[lower_(section)][lower_(name)] = to_value_(val);
      return DBS_SUCCESS;
    }

    template <typename T>
    DATABLOCK_STATUS
    replace_val(std::string const& section, std::string const& name, T const& val)
    {
      if (!has_val(section, name)) return DBS_NAME_NOT_FOUND;
      sections_[lower_(section)][lower_(name)] = to_value_(val);
      return DBS_SUCCESS;
    }

    template <typename T>
    T
    view(std::string const& section, std::string const& name) const
    {
      T out{};
      DATABLOCK_STATUS s = get_val(section, name, out);
      if (s != DBS_SUCCESS) throw BadDataBlockAccess(section, name, s);
      return out;
    }

    bool has_section(std::string const& section) const;
    bool has_val(std::string const& section, std::string const& name) const;
    // The stored value, or nullptr; *status says why not.
    value_t const* find(std::string const& section, std::string const& name,
                        DATABLOCK_STATUS* status) const;
    std::vector<std::string> section_names() const;
    std::vector<std::string> value_names(std::string const& section) const;
    void print_log(std::ostream& os) const;
    void print_log() const;
    void clear() { sections_.clear(); }

  private:
    static std::string lower_(std::string s);

    template <typename T>
    static value_t
    to_value_(T const& v)
    {
      if constexpr (std::is_same_v<T, std::vector<int>>)
        return ndarray<int>(v, {v.size()});
      else if constexpr (std::is_same_v<T, std::vector<double>>)
        return ndarray<double>(v, {v.size()});
      else if constexpr (std::is_same_v<T, char const*> || std::is_array_v<T>)
        return std::string(v);
      else
        return value_t(v);
    }

    std::map<std::string, std::map<std::string, value_t>> sections_;
  };

  template <typename T>
  DATABLOCK_STATUS
  DataBlock::get_val(std::string const& section, std::string const& name,
                     T& val) const
  {
    DATABLOCK_STATUS s;
    value_t const* v = find(section, name, &s);
    if (!v) return s;
    if constexpr (std::is_same_v<T, std::vector<int>> ||
                  std::is_same_v<T, std::vector<double>>) {
      using E = typename T::value_type;
      auto const* a = std::get_if<ndarray<E>>(v);
      if (!a || a->ndims() != 1) return DBS_WRONG_VALUE_TYPE;
      val.assign(a->begin(), a->end());
      return DBS_SUCCESS;
    } else {
      auto const* p = std::get_if<T>(v);
      if (!p) return DBS_WRONG_VALUE_TYPE;
      val = *p;
      return DBS_SUCCESS;
    }
  }

} // namespace cosmosis

#endif
