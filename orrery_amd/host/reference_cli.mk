# The reference's own command-line simulator (apps/orrery.cpp) and Python
# extension (python/src, module _orrery) with the engine's solver kinds wired in
# by reference.patch, the change a maintainer applies to the reference tree:
#
#   configuration.hpp/.cpp  SolverKind gains kHipDirect / kHipTree, spelt
#                           hip-direct / hip-tree, hip-tree a tree solver
#   assembly.cpp            make_solver hands the two kinds to the engine's
#                           make_hip_solver (orrery_hip/hip_assembly.hpp)
#   solvers/CMakeLists.txt  ORRERY_ENABLE_HIP links liborrery_hip.so
#   python sim_bindings     SolverKind.hip_direct / hip_tree
#
# This recipe stands in for that CMake option: the patched files are copies
# made under $(PATCHED) (outside this tree; the reference itself is read-only),
# every other source is compiled where it lies under REFERENCE, and the engine's
# two host sources are compiled against the reference's real headers. Outputs:
# orrery_amd/host/_cli/orrery and orrery_amd/host/_cli/_orrery<ext>, linked to
# orrery_amd/liborrery_hip.so by rpath, so both run where the reference
# checkout is absent.
#
#   make -f orrery_amd/host/reference_cli.mk -j8

REFERENCE ?= /root/reference
CLANGXX   ?= $(shell command -v clang++ 2>/dev/null || echo /opt/rocm/llvm/bin/clang++)
HERE      := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
ENGINE    := $(abspath $(HERE)..)
REPO      := $(abspath $(ENGINE)/..)
OUT       := $(HERE)_cli
OBJ       := $(OUT)/obj
PATCHED   ?= $(or $(TMPDIR),/tmp)/orrery_hip_reference_patched

VERSION   := $(shell sed -n 's/^ *VERSION \([0-9.]*\).*/\1/p' $(REFERENCE)/CMakeLists.txt | head -n1)
PYTHON    ?= python3
PY_EXT    := $(shell $(PYTHON)-config --extension-suffix)
PY_FLAGS  := $(shell $(PYTHON)-config --includes) \
             -I$(shell $(PYTHON) -c 'import pybind11; print(pybind11.get_include())')
FLAGS     := -std=c++20 -O2 -fPIC -pthread -DORRERY_ENABLE_HIP -DORRERY_HAS_AVX2_KERNEL \
             -I$(PATCHED)/include -I$(REFERENCE)/include -I$(HERE)include -I$(REPO)/include

# Every library source the simulator links, by directory under src/. The two
# patched ones come from $(PATCHED) instead.
SOURCES := core/build_info core/diagnostics core/particle_data core/random core/vec3_array \
           backend/cpu_features backend/cpu_topology backend/serial_executor \
           backend/static_executor backend/sycl_device backend/thread_pool \
           backend/work_stealing_executor \
           initial_conditions/centre_of_mass_frame initial_conditions/disc_galaxy \
           initial_conditions/galaxy_collision initial_conditions/kepler \
           initial_conditions/plummer initial_conditions/uniform_sphere \
           integrators/runge_kutta4 integrators/velocity_verlet integrators/yoshida4 \
           solvers/barnes_hut_solver solvers/direct_kernel solvers/direct_kernel_scalar \
           solvers/direct_solver solvers/morton solvers/octree solvers/reference_kernel \
           solvers/tree_walk \
           sim/binary_stream sim/checkpoint sim/config_file sim/diagnostics_log \
           sim/run_output sim/simulation sim/trajectory
PATCHED_SOURCES := sim/configuration sim/assembly
HOST_SOURCES    := hip_solvers hip_assembly
PY_SOURCES      := module array_view core_bindings initial_conditions_bindings

LIBRARY := $(addprefix $(OBJ)/,$(addsuffix .o,$(subst /,_,$(SOURCES) $(PATCHED_SOURCES)))) \
           $(OBJ)/solvers_direct_kernel_avx2.o \
           $(addprefix $(OBJ)/host_,$(addsuffix .o,$(HOST_SOURCES)))
PY_OBJECTS := $(addprefix $(OBJ)/py_,$(addsuffix .o,$(PY_SOURCES))) $(OBJ)/py_sim_bindings.o

all: $(OUT)/orrery $(OUT)/_orrery$(PY_EXT)

$(OBJ):
	mkdir -p $(OBJ)

# The patch is applied to fresh copies of every file it names, so a patch that
# no longer applies to the reference fails here rather than building stale code.
$(PATCHED)/.applied: $(HERE)reference.patch
	rm -rf $(PATCHED)
	mkdir -p $(PATCHED)/include/orrery/sim $(PATCHED)/src/sim $(PATCHED)/src/solvers $(PATCHED)/python/src
	cp $(REFERENCE)/include/orrery/sim/configuration.hpp $(PATCHED)/include/orrery/sim/
	cp $(REFERENCE)/src/sim/configuration.cpp $(REFERENCE)/src/sim/assembly.cpp $(PATCHED)/src/sim/
	cp $(REFERENCE)/src/solvers/CMakeLists.txt $(PATCHED)/src/solvers/
	cp $(REFERENCE)/python/src/sim_bindings.cpp $(PATCHED)/python/src/
	patch -p1 -s -d $(PATCHED) < $(HERE)reference.patch
	touch $@

define compile_one
$(OBJ)/$(subst /,_,$(1)).o: $(REFERENCE)/src/$(1).cpp $(PATCHED)/.applied | $(OBJ)
	$(CLANGXX) $$(FLAGS) -c $$< -o $$@
endef
$(foreach s,$(SOURCES),$(eval $(call compile_one,$(s))))

define compile_patched
$(OBJ)/$(subst /,_,$(1)).o: $(PATCHED)/.applied | $(OBJ)
	$(CLANGXX) $(FLAGS) -c $(PATCHED)/src/$(1).cpp -o $$@
endef
$(foreach s,$(PATCHED_SOURCES),$(eval $(call compile_patched,$(s))))

$(OBJ)/core_build_info.o: FLAGS += -DORRERY_VERSION_STRING='"$(VERSION)"'

$(OBJ)/solvers_direct_kernel_avx2.o: $(REFERENCE)/src/solvers/direct_kernel_avx2.cpp $(PATCHED)/.applied | $(OBJ)
	$(CLANGXX) $(FLAGS) -mavx2 -mfma -c $< -o $@

$(OBJ)/apps_orrery.o: $(REFERENCE)/apps/orrery.cpp $(PATCHED)/.applied | $(OBJ)
	$(CLANGXX) $(FLAGS) -c $< -o $@

$(OBJ)/host_%.o: $(HERE)%.cpp $(HERE)include/orrery_hip/hip_solvers.hpp \
                 $(HERE)include/orrery_hip/hip_assembly.hpp $(PATCHED)/.applied | $(OBJ)
	$(CLANGXX) $(FLAGS) -Wall -Wextra -c $< -o $@

$(OBJ)/py_%.o: $(REFERENCE)/python/src/%.cpp $(PATCHED)/.applied | $(OBJ)
	$(CLANGXX) $(FLAGS) $(PY_FLAGS) -fvisibility=hidden -I$(REFERENCE)/python/src -c $< -o $@

$(OBJ)/py_sim_bindings.o: $(PATCHED)/.applied | $(OBJ)
	$(CLANGXX) $(FLAGS) $(PY_FLAGS) -fvisibility=hidden -I$(REFERENCE)/python/src \
	    -c $(PATCHED)/python/src/sim_bindings.cpp -o $@

$(OUT)/orrery: $(LIBRARY) $(OBJ)/apps_orrery.o $(ENGINE)/liborrery_hip.so
	$(CLANGXX) -pthread -o $@.tmp $(LIBRARY) $(OBJ)/apps_orrery.o -L$(ENGINE) -lorrery_hip \
	    -Wl,-rpath,'$$ORIGIN/../..'
	mv $@.tmp $@

$(OUT)/_orrery$(PY_EXT): $(LIBRARY) $(PY_OBJECTS) $(ENGINE)/liborrery_hip.so
	$(CLANGXX) -shared -pthread -o $@.tmp $(PY_OBJECTS) $(LIBRARY) -L$(ENGINE) -lorrery_hip \
	    -Wl,-rpath,'$$ORIGIN/../..'
	mv $@.tmp $@

.PHONY: all
