# Builds oracle/_ref (test infrastructure only): the reference's own Fortran
# modules for the 2-D advection-diffusion model, compiled unmodified from
# where they lie under the reference tree (never copied into this
# repository), linked with the checker driver ref_adv2d.f90.  All outputs
# (objects, .mod files, the binary) go to oracle/_ref/.
#
#   make -f oracle/ref.mk            # -> oracle/_ref/ref_adv2d
#
# No OpenMP, as in the reference's own build/makefile (OPT=-O3, no -fopenmp):
# its !$OMP directives stay comments (Advection3DClass.f90 has some that
# flang rejects when they are read).  Module dependency order of
# src/highend/advection-diffusion2d/AdvectionClass.f90.
REF      ?= /root/reference/src
AMDFLANG ?= amdflang
FFLAGS   ?= -O2
HERE     := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT      := $(HERE)_ref

ADV2D_SRC := common/ModelPrecision common/ConstantsDictionary common/ModelFlags common/CommonRoutines \
             interp/Legendre interp/Lagrange_1D_Class interp/Lagrange_2D_Class \
             nodal/NodalStorage_2D_Class dgsem/DGSEMSolutionStorageClass_2D \
             common/LinkedListClass common/HashTableClass interp/Chebyshev interp/Lagrange_3D_Class \
             geom/CurveClass_1D geom/CurveClass_2D geom/CurveClass_3D geom/VectorClass \
             geom/GeometrySupportRoutines geom/MappedGeometryClass_2D geom/NodeClass_2D geom/EdgeClass \
             geom/QuadElementClass geom/QuadMeshClass \
             highend/advection-diffusion2d/AdvectionParamsClass highend/advection-diffusion2d/AdvectionClass
ADV2D_OBJ := $(addprefix $(OUT)/,$(addsuffix .o,$(notdir $(ADV2D_SRC))))

all: $(OUT)/ref_adv2d

# One recipe compiles the chain in order: each module needs the .mod files of
# the ones before it.
$(OUT)/ref_adv2d: $(addprefix $(REF)/,$(addsuffix .f90,$(ADV2D_SRC))) $(HERE)ref_adv2d.f90 $(HERE)ref.mk
	@mkdir -p $(OUT)
	@for f in $(ADV2D_SRC); do \
	  $(AMDFLANG) $(FFLAGS) -J$(OUT) -c $(REF)/$$f.f90 -o $(OUT)/$$(basename $$f).o || exit 1; \
	done
	$(AMDFLANG) $(FFLAGS) -J$(OUT) -c $(HERE)ref_adv2d.f90 -o $(OUT)/ref_adv2d.o
	$(AMDFLANG) $(FFLAGS) -o $@.tmp $(ADV2D_OBJ) $(OUT)/ref_adv2d.o && mv $@.tmp $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
