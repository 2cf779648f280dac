# Compiles the reference's CPU library from its sources where they lie under
# REFERENCE (read-only; nothing is copied) plus this directory's ref_shim.cpp,
# into oracle/_ref/libref.so. Test infrastructure only.
#
# clang++ because the reference's published figures and fixtures come from a
# Clang build, and Clang evaluates function arguments left to right: the
# Plummer sampler draws two unit vectors inside one argument list
# (src/initial_conditions/plummer.cpp:122), so the compiler fixes the sample.

REFERENCE ?= /root/reference
CLANGXX   ?= $(shell command -v clang++ 2>/dev/null || echo /opt/rocm/llvm/bin/clang++)
HERE      := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT       := $(HERE)_ref
FLAGS     := -std=c++20 -O2 -fPIC -ffp-contract=off -DORRERY_HAS_AVX2_KERNEL -I$(REFERENCE)/include

SOURCES := core/random core/diagnostics core/particle_data core/vec3_array \
           backend/thread_pool backend/work_stealing_executor backend/serial_executor \
           backend/static_executor backend/cpu_features backend/cpu_topology \
           initial_conditions/plummer initial_conditions/uniform_sphere \
           initial_conditions/centre_of_mass_frame integrators/velocity_verlet \
           integrators/yoshida4 integrators/runge_kutta4 solvers/direct_kernel \
           solvers/direct_kernel_scalar \
           solvers/direct_solver solvers/reference_kernel solvers/morton solvers/octree \
           solvers/tree_walk solvers/barnes_hut_solver sim/binary_stream sim/config_file \
           sim/configuration sim/checkpoint sim/trajectory

OBJECTS := $(addprefix $(OUT)/,$(addsuffix .o,$(subst /,_,$(SOURCES)))) $(OUT)/avx2.o $(OUT)/ref_shim.o

all: $(OUT)/libref.so

$(OUT):
	mkdir -p $(OUT)

define compile_one
$(OUT)/$(subst /,_,$(1)).o: $(REFERENCE)/src/$(1).cpp | $(OUT)
	$(CLANGXX) $(FLAGS) -c $$< -o $$@
endef
$(foreach s,$(SOURCES),$(eval $(call compile_one,$(s))))

$(OUT)/avx2.o: $(REFERENCE)/src/solvers/direct_kernel_avx2.cpp | $(OUT)
	$(CLANGXX) $(FLAGS) -mavx2 -mfma -c $< -o $@

$(OUT)/ref_shim.o: $(HERE)ref_shim.cpp | $(OUT)
	$(CLANGXX) $(FLAGS) -c $< -o $@

$(OUT)/libref.so: $(OBJECTS)
	$(CLANGXX) -shared -pthread -o $@ $(OBJECTS)

.PHONY: all
