# Builds the reference's native SW2D solver (sw2dSolver) straight from its
# sources under /root/reference, for use as an independent CPU cross-check of
# the oracle.  Output goes to oracle/_ref/ only (git-ignored).  Flags follow
# the reference's Release build (cmake/BuildTypes.cmake: -O3 -DNDEBUG), minus
# -static.
REF      ?= /root/reference
OUT      ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CXX      ?= g++
CXXFLAGS ?= -std=c++17 -O3 -DNDEBUG -w
SRC_COMMON := $(wildcard $(REF)/src/common/*.cpp $(REF)/src/common/geometry/*.cpp)
SRC_SOLVER := $(REF)/src/sw2dSolver/main.cpp \
              $(REF)/src/sw2dSolver/common/swSolverParam.cpp \
              $(wildcard $(REF)/src/sw2dSolver/models/*.cpp $(REF)/src/sw2dSolver/models/*/*.cpp)
OBJS := $(patsubst $(REF)/%.cpp,$(OUT)/obj/%.o,$(SRC_COMMON) $(SRC_SOLVER))

all: $(OUT)/sw2dSolver

$(OUT)/obj/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/sw2dSolver: $(OBJS)
	$(CXX) $(CXXFLAGS) $^ -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
