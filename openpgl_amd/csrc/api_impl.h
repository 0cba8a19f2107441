// Handle types and error macros shared by the C ABI translation units.
#pragma once
#include <algorithm>
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

// Host memory of a sample storage: page-locked when a HIP device is present,
// so that an update copies it to the device at the full PCIe rate (a
// pageable source goes through the runtime's staging copies at about half of
// it); ordinary memory without a device.
void* storage_alloc(size_t bytes);
void storage_free(void* p) noexcept;
template <class T>
struct StorageAllocator {
  using value_type = T;
  StorageAllocator() noexcept = default;
  template <class U>
  StorageAllocator(const StorageAllocator<U>&) noexcept {}
  T* allocate(size_t n) { return static_cast<T*>(storage_alloc(n * sizeof(T))); }
  void deallocate(T* p, size_t) noexcept { storage_free(p); }
  template <class U>
  bool operator==(const StorageAllocator<U>&) const noexcept { return true; }
  template <class U>
  bool operator!=(const StorageAllocator<U>&) const noexcept { return false; }
};
using SampleVec = std::vector<PGLSampleData, StorageAllocator<PGLSampleData>>;

// Room for n more samples: grows geometrically (a page-locked allocation is
// expensive, so batches of any size append without reallocating each time).
inline void storage_grow(SampleVec& v, size_t n) {
  if (v.size() + n > v.capacity()) v.reserve(std::max(v.size() + n, 2 * v.capacity()));
}

struct PGLSampleStorageImpl {
  SampleVec surface, volume;
  std::mutex mutex;
};

