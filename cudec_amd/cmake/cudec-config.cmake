# Package config of the MI355X cudec library: find_package(cudec 0.1.0)
# gives the static target cudec::cudec (libcudec.a: the HIP batch decoders
# for gfx950 and the C ABI of include/cudec.h, cudec_snappy.h and
# cudec_mi355x.h). Install a prefix with `python -m cudec_amd.install
# --prefix P`; this file then lies in P/lib/cmake/cudec.
#
# The archive holds HIP device code and C++ host code, so a consumer links
# the HIP runtime (libamdhip64) and the C++ runtime; both come with the
# target, so a C-only consumer links it too.

get_filename_component(_cudec_prefix "${CMAKE_CURRENT_LIST_DIR}/../../.." ABSOLUTE)

if(NOT TARGET cudec::cudec)
  find_library(CUDEC_AMDHIP64 amdhip64
               HINTS "$ENV{ROCM_PATH}/lib" /opt/rocm/lib
               NO_CACHE)
  if(NOT CUDEC_AMDHIP64)
    set(cudec_FOUND FALSE)
    set(cudec_NOT_FOUND_MESSAGE "cudec needs the HIP runtime (libamdhip64), not found")
    return()
  endif()
  include(CMakeFindDependencyMacro)
  set(THREADS_PREFER_PTHREAD_FLAG ON)
  find_dependency(Threads)

  add_library(cudec::cudec STATIC IMPORTED)
  set_target_properties(cudec::cudec PROPERTIES
    IMPORTED_LOCATION "${_cudec_prefix}/lib/libcudec.a"
    IMPORTED_LINK_INTERFACE_LANGUAGES "CXX"
    INTERFACE_INCLUDE_DIRECTORIES "${_cudec_prefix}/include"
    INTERFACE_LINK_LIBRARIES "${CUDEC_AMDHIP64};stdc++;m;Threads::Threads;${CMAKE_DL_LIBS}")
endif()
unset(_cudec_prefix)
