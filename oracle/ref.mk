# Builds the reference NebulaSEM CPU path from its own sources, in place under
# /root/reference, into oracle/_ref/ (never committed).  This is the checker's
# reference build: test infrastructure only, never linked into the product.
#
#   make -f oracle/ref.mk -j8          (from the repository root)
#
# The reference's own build system is not used.  Its sources need MPI; the
# image's MPICH (/opt/conda) provides it.  No stand-ins are written for
# anything.  Compiled without -march=native and with -ffp-contract=off so the
# floating-point results do not depend on the host CPU's FMA support.

REF      ?= /root/reference
OUT      ?= oracle/_ref
OBJ      ?= oracle/_obj
MPI_INC  ?= /opt/conda/include
MPI_LIB  ?= /opt/conda/lib
CXX      := g++
CXXFLAGS := -std=c++17 -O2 -ffp-contract=off -DUSE_DOUBLE -DUSE_EXPR_TMPL -w \
            -I$(MPI_INC) -I$(REF)/src/field -I$(REF)/src/mesh -I$(REF)/src/mp \
            -I$(REF)/src/prepare -I$(REF)/src/solvers -I$(REF)/src/tensor \
            -I$(REF)/src/util -I$(REF)/src/vtk -I$(REF)/src/turbulence \
            -I$(REF)/src/turbulence/ke -I$(REF)/src/turbulence/kw \
            -I$(REF)/src/turbulence/les -I$(REF)/src/turbulence/mixing_length \
            -I$(REF)/src/turbulence/realizableke -I$(REF)/src/turbulence/rngke \
            -I$(REF)/apps/utils
LDFLAGS  := $(MPI_LIB)/libmpicxx.so $(MPI_LIB)/libmpi.so -Wl,-rpath,/usr/lib/x86_64-linux-gnu:$(MPI_LIB)

LIBSRC := src/field/dg.cpp src/field/field.cpp src/mesh/mesh.cpp \
          src/mesh/hexMesh.cpp src/mesh/mshMesh.cpp src/mp/mp.cpp \
          src/prepare/prepare.cpp src/solvers/solve.cpp src/tensor/tensor.cpp \
          src/util/util.cpp src/vtk/vtk.cpp src/turbulence/turbulence.cpp \
          src/turbulence/ke/ke.cpp src/turbulence/kw/kw.cpp \
          src/turbulence/les/les.cpp src/turbulence/mixing_length/mixing_length.cpp \
          src/turbulence/realizableke/realizableke.cpp src/turbulence/rngke/rngke.cpp \
          apps/utils/properties.cpp apps/utils/wrapper.cpp apps/utils/calc_walldist.cpp
LIBOBJ := $(patsubst %.cpp,$(OBJ)/%.o,$(LIBSRC))

APPS := mesh euler potential piso prepare
APPSRC_mesh      := apps/mesh/meshApp.cpp
APPSRC_euler     := apps/euler/euler.cpp
APPSRC_potential := apps/potential/potential.cpp
APPSRC_piso      := apps/piso/piso.cpp
APPSRC_prepare   := apps/prepare/prepareApp.cpp

all: $(addprefix $(OUT)/,$(APPS)) $(OUT)/geomdump $(OUT)/euler_tendency

$(OBJ)/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/libnebulasem_ref.a: $(LIBOBJ)
	@mkdir -p $(OUT)
	ar rcs $@ $^

define APP_RULE
$(OUT)/$(1): $(OBJ)/$(patsubst %.cpp,%.o,$(APPSRC_$(1))) $(OUT)/libnebulasem_ref.a
	$(CXX) -o $$@ $$^ $(LDFLAGS)
endef
$(foreach a,$(APPS),$(eval $(call APP_RULE,$(a))))

# geometry dump driver: this repository's own code, linked against the
# reference library, writes the reference's loaded-mesh arrays for tests.
$(OBJ)/geomdump.o: oracle/ref_geomdump.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/geomdump: $(OBJ)/geomdump.o $(OUT)/libnebulasem_ref.a
	$(CXX) -o $@ $^ $(LDFLAGS)

# tendency dump driver: this repository's own code, linked against the
# reference library, evaluates the reference's divergence operators of the
# Euler fluxes at a given state (the method-of-lines residual's pin).
$(OBJ)/euler_tendency.o: oracle/ref_tendency.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS) -c $< -o $@

$(OUT)/euler_tendency: $(OBJ)/euler_tendency.o $(OUT)/libnebulasem_ref.a
	$(CXX) -o $@ $^ $(LDFLAGS)

# The reference's OpenMP build of the Euler and Poisson apps (its USE_OMP
# option, -O3 as its release build), for timing the CPU baseline on the host
# cores; results are not compared bit-for-bit with anything.
OBJOMP   ?= oracle/_obj_omp
OMPFLAGS := $(filter-out -O2 -ffp-contract=off,$(CXXFLAGS)) -O3 -fopenmp
LIBOBJOMP := $(patsubst %.cpp,$(OBJOMP)/%.o,$(LIBSRC))

$(OBJOMP)/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(OMPFLAGS) -c $< -o $@

$(OUT)/libnebulasem_ref_omp.a: $(LIBOBJOMP)
	@mkdir -p $(OUT)
	ar rcs $@ $^

$(OUT)/euler_omp: $(OBJOMP)/apps/euler/euler.o $(OUT)/libnebulasem_ref_omp.a
	$(CXX) -fopenmp -o $@ $^ $(LDFLAGS)

$(OUT)/potential_omp: $(OBJOMP)/apps/potential/potential.o $(OUT)/libnebulasem_ref_omp.a
	$(CXX) -fopenmp -o $@ $^ $(LDFLAGS)

all: $(OUT)/euler_omp $(OUT)/potential_omp

.PHONY: all
