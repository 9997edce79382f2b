# Builds the reference's whole renderer (core/, base/ and extension/ with the
# vendored miniz and zstd decoder) from the reference sources where they lie,
# plus oracle/ref_render_driver.cpp, into oracle/_ref/sprout_render (never
# committed).  base/platform/platform.cpp is left out: it includes a version
# header that the reference's CMake build generates, and nothing on the
# rendering path calls it; so are core/testing/ (developer self-tests).
# Flags as oracle/ref.mk (the reference's Release
# configuration on Linux).
#
#   make -f oracle/ref_render.mk -j8                  # native build (golden renders)
#   make -f oracle/ref_render.mk -j8 ARCH=x86-64-v3   # portable build (CPU baseline on other hosts)
#
# Only this container has the reference; the render fixtures it makes are
# committed under tests/golden/ (tests/golden/make_sprout_renders.py).

REF   ?= /root/reference
ARCH  ?= native
HERE  := $(dir $(lastword $(MAKEFILE_LIST)))
OUT   := $(HERE)_ref
SRC   := $(REF)/source

ifeq ($(ARCH),native)
OBJ   := $(OUT)/render_obj
BIN   := $(OUT)/sprout_render
else
OBJ   := $(OUT)/render_obj_$(subst -,_,$(ARCH))
BIN   := $(OUT)/sprout_render_$(subst -,_,$(ARCH))
endif

CXXFLAGS := -std=c++20 -O3 -DNDEBUG -march=$(ARCH) -fno-rtti -fno-exceptions \
            -DSU_ACESCG -DSU_IBL_MIS_COMPENSATION -DSU_ZSTD -w \
            -I$(SRC)/core -I$(SRC) -I$(SRC)/base -I$(REF)/thirdparty/include

CPP  := $(filter-out $(SRC)/base/platform/platform.cpp $(SRC)/core/testing/%, \
          $(shell find $(SRC)/core $(SRC)/base $(SRC)/extension -name '*.cpp'))
OBJS := $(patsubst $(SRC)/%.cpp,$(OBJ)/%.o,$(CPP)) \
        $(OBJ)/thirdparty/miniz.o $(OBJ)/thirdparty/zstddeclib.o

all: $(BIN)

$(OBJ)/%.o: $(SRC)/%.cpp
	@mkdir -p $(dir $@)
	g++ $(CXXFLAGS) -c $< -o $@

$(OBJ)/thirdparty/miniz.o: $(REF)/thirdparty/include/miniz/miniz.cpp
	@mkdir -p $(dir $@)
	g++ $(CXXFLAGS) -c $< -o $@

$(OBJ)/thirdparty/zstddeclib.o: $(REF)/thirdparty/include/zstd/zstddeclib.cpp
	@mkdir -p $(dir $@)
	g++ $(CXXFLAGS) -c $< -o $@

$(BIN): $(OBJS) $(HERE)ref_render_driver.cpp
	g++ $(CXXFLAGS) -o $@ $(HERE)ref_render_driver.cpp $(OBJS) -lpthread

.PHONY: all
