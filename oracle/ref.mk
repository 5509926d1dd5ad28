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

# _E physics of the 4D-Var layer (the files that do not need the Enzyme
# runtime) driven by ref_e_driver.cpp through the dry-run loop
SRC_E := $(REF)/assim/swSpE_hyp.cpp $(REF)/assim/swSpE_friction.cpp $(REF)/assim/swSpE_core.cpp \
         $(REF)/assim/swSpE_muscl.cpp
OBJS_E := $(patsubst $(REF)/%.cpp,$(OUT)/obj/%.o,$(SRC_E))
OBJS_NOMAIN := $(filter-out $(OUT)/obj/src/sw2dSolver/main.o,$(OBJS))
EINC := -I$(REF)/assim -I$(REF)/src/sw2dSolver/models/sp -I$(REF)/src/sw2dSolver/models

$(OUT)/obj/assim/%.o: $(REF)/assim/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -ffunction-sections -fdata-sections $(EINC) -c $< -o $@

$(OUT)/ref_e_driver: $(dir $(lastword $(MAKEFILE_LIST)))ref_e_driver.cpp $(OBJS_E) $(OBJS_NOMAIN)
	$(CXX) $(CXXFLAGS) $(EINC) -I$(REF)/assim $< $(OBJS_E) $(OBJS_NOMAIN) -Wl,--gc-sections -o $@

all: $(OUT)/ref_e_driver
