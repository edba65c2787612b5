/*
 * c_datablock.h -- the C API over cosmosis::DataBlock (subset of CosmoSIS's
 * c_datablock.h with the same function names and argument meanings).  A
 * c_datablock* is the cosmosis::DataBlock* that CosmoSIS passes to a
 * module's setup() and execute().
 */
#ifndef Y3G_SHIM_C_DATABLOCK_H
#define Y3G_SHIM_C_DATABLOCK_H

#include "cosmosis/datablock/datablock_status.h"
#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void c_datablock;

c_datablock* make_c_datablock(void);
DATABLOCK_STATUS destroy_c_datablock(c_datablock* s);

bool c_datablock_has_section(c_datablock const* s, const char* section);
bool c_datablock_has_value(c_datablock const* s, const char* section,
                           const char* name);

DATABLOCK_STATUS c_datablock_put_int(c_datablock* s, const char* section,
                                     const char* name, int val);
DATABLOCK_STATUS c_datablock_put_double(c_datablock* s, const char* section,
                                        const char* name, double val);
DATABLOCK_STATUS c_datablock_put_bool(c_datablock* s, const char* section,
                                      const char* name, bool val);
DATABLOCK_STATUS c_datablock_put_string(c_datablock* s, const char* section,
                                        const char* name, const char* val);
DATABLOCK_STATUS c_datablock_put_int_array_1d(c_datablock* s,
                                              const char* section,
                                              const char* name,
                                              int const* val, int size);
DATABLOCK_STATUS c_datablock_put_double_array_1d(c_datablock* s,
                                                 const char* section,
                                                 const char* name,
                                                 double const* val, int size);
DATABLOCK_STATUS c_datablock_put_double_array(c_datablock* s,
                                              const char* section,
                                              const char* name,
                                              double const* val, int ndims,
                                              int const* extents);

DATABLOCK_STATUS c_datablock_get_int(c_datablock* s, const char* section,
                                     const char* name, int* val);
DATABLOCK_STATUS c_datablock_get_double(c_datablock* s, const char* section,
                                        const char* name, double* val);
DATABLOCK_STATUS c_datablock_get_bool(c_datablock* s, const char* section,
                                      const char* name, bool* val);

/* Length of a 1-d array, or -1. */
int c_datablock_get_array_length(c_datablock* s, const char* section,
                                 const char* name);
DATABLOCK_STATUS c_datablock_get_array_ndim(c_datablock* s,
                                            const char* section,
                                            const char* name, int* ndim);
DATABLOCK_STATUS c_datablock_get_int_array_1d_preallocated(
  c_datablock* s, const char* section, const char* name, int* array,
  int* size, int maxsize);
DATABLOCK_STATUS c_datablock_get_double_array_1d_preallocated(
  c_datablock* s, const char* section, const char* name, double* array,
  int* size, int maxsize);
DATABLOCK_STATUS c_datablock_get_double_array_shape(c_datablock* s,
                                                    const char* section,
                                                    const char* name,
                                                    int ndims, int* extents);
DATABLOCK_STATUS c_datablock_get_int_array_shape(c_datablock* s,
                                                 const char* section,
                                                 const char* name, int ndims,
                                                 int* extents);
DATABLOCK_STATUS c_datablock_get_double_array(c_datablock* s,
                                              const char* section,
                                              const char* name, double* val,
                                              int ndims, int const* extents);
DATABLOCK_STATUS c_datablock_get_int_array(c_datablock* s, const char* section,
                                           const char* name, int* val,
                                           int ndims, int const* extents);

#ifdef __cplusplus
}
#endif
#endif
