// datablock_shim.cc -- the non-template members of cosmosis::DataBlock and
// the C API of c_datablock.h (libdatablock.so).
#include "cosmosis/datablock/c_datablock.h"
#include "cosmosis/datablock/datablock.hh"

#include <algorithm>
#include <cctype>
#include <iostream>

namespace cosmosis {

  std::string
  DataBlock::lower_(std::string s)
  {
    std::transform(s.begin(), s.end(), s.begin(),
                   [](unsigned char c) { return std::tolower(c); });
    return s;
  }

  bool
  DataBlock::has_section(std::string const& section) const
  {
    return sections_.count(lower_(section)) != 0;
  }

  bool
  DataBlock::has_val(std::string const& section, std::string const& name) const
  {
    DATABLOCK_STATUS s;
    return find(section, name, &s) != nullptr;
  }

  DataBlock::value_t const*
  DataBlock::find(std::string const& section, std::string const& name,
                  DATABLOCK_STATUS* status) const
  {
    auto is = sections_.find(lower_(section));
    if (is == sections_.end()) {
      *status = DBS_SECTION_NOT_FOUND;
      return nullptr;
    }
    auto iv = is->second.find(lower_(name));
    if (iv == is->second.end()) {
      *status = DBS_NAME_NOT_FOUND;
      return nullptr;
    }
    *status = DBS_SUCCESS;
    return &iv->second;
  }

  std::vector<std::string>
  DataBlock::section_names() const
  {
    std::vector<std::string> out;
    for (auto const& kv : sections_) out.push_back(kv.first);
    return out;
  }

  std::vector<std::string>
  DataBlock::value_names(std::string const& section) const
  {
    std::vector<std::string> out;
    auto is = sections_.find(lower_(section));
    if (is != sections_.end())
      for (auto const& kv : is->second) out.push_back(kv.first);
    return out;
  }

  void
  DataBlock::print_log(std::ostream& os) const
  {
    for (auto const& s : sections_)
      for (auto const& v : s.second)
        os << s.first << '/' << v.first << " (type index " << v.second.index()
           << ")\n";
  }

  void
  DataBlock::print_log() const
  {
    print_log(std::cerr);
  }

} // namespace cosmosis

using cosmosis::DataBlock;
using cosmosis::ndarray;

namespace {

  DataBlock*
  db(c_datablock* s)
  {
    return static_cast<DataBlock*>(s);
  }
  DataBlock const*
  db(c_datablock const* s)
  {
    return static_cast<DataBlock const*>(s);
  }

#define Y3G_DB_CHECK(s, section, name)                                        \
  do {                                                                         \
    if (!(s)) return DBS_DATABLOCK_NULL;                                       \
    if (!(section)) return DBS_SECTION_NULL;                                   \
    if (!(name)) return DBS_NAME_NULL;                                         \
  } while (0)

  template <typename T>
  DATABLOCK_STATUS
  put(c_datablock* s, const char* section, const char* name, T const& v)
  {
    Y3G_DB_CHECK(s, section, name);
    try {
      return db(s)->put_val(section, name, v);
    }
    catch (...) {
      return DBS_LOGIC_ERROR;
    }
  }

  template <typename T>
  DATABLOCK_STATUS
  get(c_datablock* s, const char* section, const char* name, T* out)
  {
    Y3G_DB_CHECK(s, section, name);
    if (!out) return DBS_VALUE_NULL;
    return db(s)->get_val(section, name, *out);
  }

  template <typename T>
  DATABLOCK_STATUS
  get_array(c_datablock* s, const char* section, const char* name,
            ndarray<T> const** out)
  {
    Y3G_DB_CHECK(s, section, name);
    DATABLOCK_STATUS st;
    auto const* v = db(s)->find(section, name, &st);
    if (!v) return st;
    auto const* a = std::get_if<ndarray<T>>(v);
    if (!a) return DBS_WRONG_VALUE_TYPE;
    *out = a;
    return DBS_SUCCESS;
  }

  template <typename T>
  DATABLOCK_STATUS
  get_1d_pre(c_datablock* s, const char* section, const char* name, T* array,
             int* size, int maxsize)
  {
    ndarray<T> const* a = nullptr;
    DATABLOCK_STATUS st = get_array(s, section, name, &a);
    if (st != DBS_SUCCESS) return st;
    if (!array || !size) return DBS_VALUE_NULL;
    if (a->ndims() != 1) return DBS_NDIM_MISMATCH;
    *size = static_cast<int>(a->size());
    if (static_cast<int>(a->size()) > maxsize) return DBS_SIZE_INSUFFICIENT;
    std::copy(a->begin(), a->end(), array);
    return DBS_SUCCESS;
  }

  template <typename T>
  DATABLOCK_STATUS
  get_shape(c_datablock* s, const char* section, const char* name, int ndims,
            int* extents)
  {
    ndarray<T> const* a = nullptr;
    DATABLOCK_STATUS st = get_array(s, section, name, &a);
    if (st != DBS_SUCCESS) return st;
    if (!extents) return DBS_EXTENTS_NULL;
    if (ndims != static_cast<int>(a->ndims())) return DBS_NDIM_MISMATCH;
    for (int i = 0; i != ndims; ++i) extents[i] = static_cast<int>(a->extents()[i]);
    return DBS_SUCCESS;
  }

  template <typename T>
  DATABLOCK_STATUS
  get_nd(c_datablock* s, const char* section, const char* name, T* val,
         int ndims, int const* extents)
  {
    ndarray<T> const* a = nullptr;
    DATABLOCK_STATUS st = get_array(s, section, name, &a);
    if (st != DBS_SUCCESS) return st;
    if (!val) return DBS_VALUE_NULL;
    if (!extents) return DBS_EXTENTS_NULL;
    if (ndims != static_cast<int>(a->ndims())) return DBS_NDIM_MISMATCH;
    for (int i = 0; i != ndims; ++i)
      if (extents[i] != static_cast<int>(a->extents()[i])) return DBS_EXTENTS_MISMATCH;
    std::copy(a->begin(), a->end(), val);
    return DBS_SUCCESS;
  }

} // namespace

extern "C" {

c_datablock*
make_c_datablock(void)
{
  return new DataBlock;
}

DATABLOCK_STATUS
destroy_c_datablock(c_datablock* s)
{
  if (!s) return DBS_DATABLOCK_NULL;
  delete db(s);
  return DBS_SUCCESS;
}

bool
c_datablock_has_section(c_datablock const* s, const char* section)
{
  return s && section && db(s)->has_section(section);
}

bool
c_datablock_has_value(c_datablock const* s, const char* section, const char* name)
{
  return s && section && name && db(s)->has_val(section, name);
}

DATABLOCK_STATUS
c_datablock_put_int(c_datablock* s, const char* section, const char* name, int val)
{
  return put(s, section, name, val);
}

DATABLOCK_STATUS
c_datablock_put_double(c_datablock* s, const char* section, const char* name,
                       double val)
{
  return put(s, section, name, val);
}

DATABLOCK_STATUS
c_datablock_put_bool(c_datablock* s, const char* section, const char* name,
                     bool val)
{
  return put(s, section, name, val);
}

DATABLOCK_STATUS
c_datablock_put_string(c_datablock* s, const char* section, const char* name,
                       const char* val)
{
  if (!val) return DBS_VALUE_NULL;
  return put(s, section, name, std::string(val));
}

DATABLOCK_STATUS
c_datablock_put_int_array_1d(c_datablock* s, const char* section,
                             const char* name, int const* val, int size)
{
  if (!val && size > 0) return DBS_VALUE_NULL;
  if (size < 0) return DBS_SIZE_NONPOSITIVE;
  return put(s, section, name, std::vector<int>(val, val + size));
}

DATABLOCK_STATUS
c_datablock_put_double_array_1d(c_datablock* s, const char* section,
                                const char* name, double const* val, int size)
{
  if (!val && size > 0) return DBS_VALUE_NULL;
  if (size < 0) return DBS_SIZE_NONPOSITIVE;
  return put(s, section, name, std::vector<double>(val, val + size));
}

DATABLOCK_STATUS
c_datablock_put_double_array(c_datablock* s, const char* section,
                             const char* name, double const* val, int ndims,
                             int const* extents)
{
  if (!val) return DBS_VALUE_NULL;
  if (!extents) return DBS_EXTENTS_NULL;
  if (ndims <= 0) return DBS_NDIM_NONPOSITIVE;
  try {
    return put(s, section, name, ndarray<double>(val, ndims, extents));
  }
  catch (...) {
    return DBS_EXTENTS_MISMATCH;
  }
}

DATABLOCK_STATUS
c_datablock_get_int(c_datablock* s, const char* section, const char* name, int* val)
{
  return get(s, section, name, val);
}

DATABLOCK_STATUS
c_datablock_get_double(c_datablock* s, const char* section, const char* name,
                       double* val)
{
  return get(s, section, name, val);
}

DATABLOCK_STATUS
c_datablock_get_bool(c_datablock* s, const char* section, const char* name,
                     bool* val)
{
  return get(s, section, name, val);
}

int
c_datablock_get_array_length(c_datablock* s, const char* section, const char* name)
{
  if (!s || !section || !name) return -1;
  DATABLOCK_STATUS st;
  auto const* v = db(s)->find(section, name, &st);
  if (!v) return -1;
  if (auto const* a = std::get_if<ndarray<double>>(v))
    return a->ndims() == 1 ? static_cast<int>(a->size()) : -1;
  if (auto const* a = std::get_if<ndarray<int>>(v))
    return a->ndims() == 1 ? static_cast<int>(a->size()) : -1;
  if (auto const* a = std::get_if<std::vector<std::string>>(v))
    return static_cast<int>(a->size());
  return -1;
}

DATABLOCK_STATUS
c_datablock_get_array_ndim(c_datablock* s, const char* section, const char* name,
                           int* ndim)
{
  Y3G_DB_CHECK(s, section, name);
  if (!ndim) return DBS_VALUE_NULL;
  DATABLOCK_STATUS st;
  auto const* v = db(s)->find(section, name, &st);
  if (!v) return st;
  if (auto const* a = std::get_if<ndarray<double>>(v))
    *ndim = static_cast<int>(a->ndims());
  else if (auto const* a = std::get_if<ndarray<int>>(v))
    *ndim = static_cast<int>(a->ndims());
  else
    return DBS_WRONG_VALUE_TYPE;
  return DBS_SUCCESS;
}

DATABLOCK_STATUS
c_datablock_get_int_array_1d_preallocated(c_datablock* s, const char* section,
                                          const char* name, int* array,
                                          int* size, int maxsize)
{
  return get_1d_pre(s, section, name, array, size, maxsize);
}

DATABLOCK_STATUS
c_datablock_get_double_array_1d_preallocated(c_datablock* s, const char* section,
                                             const char* name, double* array,
                                             int* size, int maxsize)
{
  return get_1d_pre(s, section, name, array, size, maxsize);
}

DATABLOCK_STATUS
c_datablock_get_double_array_shape(c_datablock* s, const char* section,
                                   const char* name, int ndims, int* extents)
{
  return get_shape<double>(s, section, name, ndims, extents);
}

DATABLOCK_STATUS
c_datablock_get_int_array_shape(c_datablock* s, const char* section,
                                const char* name, int ndims, int* extents)
{
  return get_shape<int>(s, section, name, ndims, extents);
}

DATABLOCK_STATUS
c_datablock_get_double_array(c_datablock* s, const char* section,
                             const char* name, double* val, int ndims,
                             int const* extents)
{
  return get_nd(s, section, name, val, ndims, extents);
}

DATABLOCK_STATUS
c_datablock_get_int_array(c_datablock* s, const char* section, const char* name,
                          int* val, int ndims, int const* extents)
{
  return get_nd(s, section, name, val, ndims, extents);
}

} // extern "C"
