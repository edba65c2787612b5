/*
 * datablock_status.h -- status codes of the CosmoSIS DataBlock API.
 *
 * Part of a source-compatible subset of CosmoSIS's cosmosis/datablock
 * headers, enough to build and test the integration modules without a
 * CosmoSIS installation.  Against a real CosmoSIS, put its include directory
 * first on the include path instead (modules/Makefile: COSMOSIS_SRC_DIR).
 * Only DBS_SUCCESS == 0 is relied upon by the modules.
 */
#ifndef Y3G_SHIM_DATABLOCK_STATUS_H
#define Y3G_SHIM_DATABLOCK_STATUS_H

typedef enum {
  DBS_SUCCESS = 0,
  DBS_DATABLOCK_NULL,
  DBS_SECTION_NULL,
  DBS_SECTION_NOT_FOUND,
  DBS_NAME_NULL,
  DBS_NAME_NOT_FOUND,
  DBS_NAME_ALREADY_EXISTS,
  DBS_VALUE_NULL,
  DBS_WRONG_VALUE_TYPE,
  DBS_MEMORY_ALLOC_FAILURE,
  DBS_SIZE_NULL,
  DBS_SIZE_NONPOSITIVE,
  DBS_SIZE_INSUFFICIENT,
  DBS_NDIM_NONPOSITIVE,
  DBS_NDIM_OVERFLOW,
  DBS_NDIM_MISMATCH,
  DBS_EXTENTS_NULL,
  DBS_EXTENTS_MISMATCH,
  DBS_LOGIC_ERROR,
  DBS_USED_DEFAULT
} DATABLOCK_STATUS;

#endif
