// Handle types and error macros shared by the C ABI translation units.
#pragma once
#include <iostream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/openpgl_amd/openpgl.h"
#include "field.h"
#include "path_batch.h"

// Error convention of the reference's C API (api/api.cpp:33-72): creation,
// update, reset, release and file functions print the exception and return
// nullptr / false / nothing.
#define PGL_CATCH_BEGIN try {
#define PGL_CATCH_END(ret)                                            \
  }                                                                   \
  catch (const std::bad_alloc&) {                                     \
    std::cout << "Open PGL was unable to allocate memory" << std::endl; \
    return ret;                                                       \
  }                                                                   \
  catch (const std::exception& e) {                                   \
    std::cout << e.what() << std::endl;                               \
    return ret;                                                       \
  }                                                                   \
  catch (...) {                                                       \
    std::cout << "an unrecognized exception was caught" << std::endl; \
    return ret;                                                       \
  }
#define PGL_CATCH_END_VOID                                            \
  }                                                                   \
  catch (const std::bad_alloc&) {                                     \
    std::cout << "Open PGL was unable to allocate memory" << std::endl; \
    return;                                                           \
  }                                                                   \
  catch (const std::exception& e) {                                   \
    std::cout << e.what() << std::endl;                               \
    return;                                                           \
  }                                                                   \
  catch (...) {                                                       \
    std::cout << "an unrecognized exception was caught" << std::endl; \
    return;                                                           \
  }
#define THROW_IF_NULL(obj, name) \
  if (obj == nullptr) throw std::runtime_error(std::string("null ") + name + " provided to " + __FUNCTION__)

struct PGLDeviceImpl {
  int type;
  int vectorWidth;
};

struct PGLFieldImpl {
  pglgpu::GuidingField* gf;
  int deviceType;
  std::vector<PGLSampleData>* staging = nullptr;
  pglgpu::DevBuf dStage;
  pglgpu::PathSampleBatch* paths = nullptr;  // path-segment updates (created on first use)
};

struct PGLSampleStorageImpl {
  std::vector<PGLSampleData> surface, volume;
  std::mutex mutex;
};

