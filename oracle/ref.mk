# Builds the reference's own triangle SBVH builder and traversal, its
# Substitute (Disney + GGX) material sample, its sensor filters, its random
# generator, Golden_ratio camera sampler and sRGB -> AP1 mapping, and its
# JSON mesh handler and binary mesh writer, from the reference sources where
# they lie, into oracle/_ref/ (never committed).
# Flags follow the reference's CMake configuration for a Release build on
# Linux (source/CMakeLists.txt: C++20, -O3, -march=native, no RTTI, no
# exceptions, SU_ACESCG / SU_IBL_MIS_COMPENSATION / SU_ZSTD definitions).
#
#   make -f oracle/ref.mk                 # native build (golden vectors)
#   make -f oracle/ref.mk ARCH=x86-64-v3  # portable build (CPU baseline on other hosts)
#
# Only this container has the reference; the built .so travels with the tree.

REF   ?= /root/reference
ARCH  ?= native
HERE  := $(dir $(lastword $(MAKEFILE_LIST)))
OUT   := $(HERE)_ref
SRC   := $(REF)/source

ifeq ($(ARCH),native)
LIB   := $(OUT)/libsprout_ref.so
else
LIB   := $(OUT)/libsprout_ref_$(subst -,_,$(ARCH)).so
endif

CXXFLAGS := -std=c++20 -O3 -DNDEBUG -march=$(ARCH) -fno-rtti -fno-exceptions -fPIC \
            -DSU_ACESCG -DSU_IBL_MIS_COMPENSATION -DSU_ZSTD -w \
            -I$(SRC)/core -I$(SRC) -I$(SRC)/base -I$(REF)/thirdparty/include

REF_SOURCES := \
  $(SRC)/core/scene/bvh/scene_bvh_builder_base.cpp \
  $(SRC)/core/scene/bvh/scene_bvh_builder.cpp \
  $(SRC)/core/scene/shape/triangle/bvh/triangle_bvh_builder_sah.cpp \
  $(SRC)/core/scene/shape/triangle/bvh/triangle_bvh_helper.cpp \
  $(SRC)/core/scene/shape/triangle/bvh/triangle_bvh_tree.cpp \
  $(SRC)/core/scene/shape/shape_vertex.cpp \
  $(SRC)/base/thread/thread_pool.cpp \
  $(SRC)/core/logging/logging.cpp \
  $(SRC)/core/logging/log.cpp \
  $(SRC)/core/scene/material/substitute/substitute_sample.cpp \
  $(SRC)/core/sampler/sampler.cpp \
  $(SRC)/core/sampler/sampler_golden_ratio.cpp \
  $(SRC)/core/rendering/sensor/filter/sensor_gaussian.cpp \
  $(SRC)/core/rendering/sensor/filter/sensor_blackman.cpp \
  $(SRC)/core/rendering/sensor/filter/sensor_mitchell.cpp \
  $(SRC)/core/scene/shape/triangle/triangle_json_handler.cpp \
  $(SRC)/core/scene/shape/triangle/triangle_mesh_exporter.cpp \
  $(SRC)/core/scene/shape/triangle/vertex_layout_description.cpp \
  $(SRC)/base/string/string.cpp

all: $(LIB)

DRIVERS := $(HERE)ref_driver.cpp $(HERE)ref_bsdf_driver.cpp $(HERE)ref_filter_driver.cpp \
           $(HERE)ref_sampler_driver.cpp $(HERE)ref_mesh_driver.cpp

$(LIB): $(REF_SOURCES) $(DRIVERS)
	@mkdir -p $(OUT)
	g++ $(CXXFLAGS) -shared -o $@ $(DRIVERS) $(REF_SOURCES) -lpthread

.PHONY: all
