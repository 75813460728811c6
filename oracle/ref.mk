# Builds oracle/_ref (test infrastructure only): the reference's own Fortran
# modules for the 2-D advection-diffusion model, the 3-D advection model and
# the 2-D CG spectral-element elliptic solver, and its Krylov loops, compiled
# unmodified from where they lie under the reference tree (never copied into
# this repository), linked with the checker drivers ref_adv2d.f90,
# ref_adv3d.f90, ref_cgsem2d.f90 and ref_krylov.f90, the quadrature,
# nodal-storage and interpolation modules behind ref_nodal.f90, the 2-D
# advection model's pickup routines behind ref_pickup.f90, and the SpecMesh
# reader and element geometry behind ref_specmesh.f90, and the conservative
# shallow-water class of the planewave example behind ref_sw2d.f90.  All outputs
# (objects, .mod files, binaries) go to oracle/_ref/.
#
#   make -f oracle/ref.mk   # -> oracle/_ref/ref_{adv2d,adv3d,cgsem2d,krylov,nodal,pickup,specmesh,sw2d}, the
#                           # reference's example programs (ref_*) and the same programs on the drop-in (dropin_*)
#
# No OpenMP, as in the reference's own build/makefile (OPT=-O3, no -fopenmp):
# its !$OMP directives stay comments (Advection3DClass.f90 has some that
# flang rejects when they are read).  The two models each define a module
# named AdvectionParamsClass, so each model's own modules get a directory of
# their own; the shared modules live in oracle/_ref.
REF      ?= /root/reference/src
AMDFLANG ?= amdflang
FFLAGS   ?= -O2
HERE     := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT      := $(HERE)_ref

# Shared modules in dependency order.
COMMON_SRC := common/ModelPrecision common/ConstantsDictionary common/ModelFlags common/CommonRoutines \
              common/Timing common/LinkedListClass common/KeyRingClass common/HashTableClass \
              interp/Legendre interp/Chebyshev interp/Lagrange_1D_Class interp/Lagrange_2D_Class \
              interp/Lagrange_3D_Class \
              nodal/NodalStorage_2D_Class nodal/NodalStorage_3D_Class \
              dgsem/DGSEMSolutionStorageClass_2D dgsem/DGSEMSolutionStorageClass_3D \
              filters/RollOffFilter3D_Class \
              geom/CurveClass_1D geom/CurveClass_2D geom/CurveClass_3D geom/VectorClass \
              geom/GeometrySupportRoutines geom/MappedGeometryClass_2D geom/NodeClass_2D geom/EdgeClass \
              geom/QuadElementClass geom/QuadMeshClass \
              geom/SurfaceClass_3D geom/MappedGeometryClass_3D geom/NodeClass_3D geom/FaceClass \
              geom/HexElementClass geom/HexMeshClass
COMMON_OBJ := $(addprefix $(OUT)/,$(addsuffix .o,$(notdir $(COMMON_SRC))))

ADV2D_SRC := highend/advection-diffusion2d/AdvectionParamsClass highend/advection-diffusion2d/AdvectionClass
ADV3D_SRC := highend/advection3d/Advection3DParamsClass highend/advection3d/Advection3DClass
CGSEM2D_SRC := cgsem/CGsemParams_Class cgsem/CGsemPreconditioner_2D_Class cgsem/CGsemElliptic_2D_Class
KRYLOV_SRC := iterativesolve/IterativeSolversParams_Class iterativesolve/IterativeSolvers_Class
# The conservative shallow-water class as the planewave example ships it (the
# src/highend/shallowwater copy does not compile); paths relative to $(REF).
SW2D_SRC := ../examples/shallowwater/planewave/localmods/SWParamsClass \
            ../examples/shallowwater/planewave/localmods/ConservativeShallowWaterClass

all: $(OUT)/ref_adv2d $(OUT)/ref_adv3d $(OUT)/ref_cgsem2d $(OUT)/ref_krylov $(OUT)/ref_nodal $(OUT)/ref_pickup \
     $(OUT)/ref_specmesh $(OUT)/ref_sw2d $(OUT)/ref_cgsem_paraboloid $(OUT)/dropin_cgsem_paraboloid \
     $(OUT)/ref_adv2d_ic $(OUT)/ref_adv2d_driver $(OUT)/dropin_adv2d_driver \
     $(OUT)/ref_adv3d_ic $(OUT)/ref_adv3d_driver $(OUT)/dropin_adv3d_ic $(OUT)/dropin_adv3d_driver \
     $(OUT)/dropin_elliptic1d_driver

# One recipe compiles the shared chain in order: each module needs the .mod
# files of the ones before it.
$(OUT)/common.stamp: $(addprefix $(REF)/,$(addsuffix .f90,$(COMMON_SRC))) $(HERE)ref.mk
	@mkdir -p $(OUT)
	@for f in $(COMMON_SRC); do \
	  $(AMDFLANG) $(FFLAGS) -J$(OUT) -c $(REF)/$$f.f90 -o $(OUT)/$$(basename $$f).o || exit 1; \
	done
	@touch $@

# $(call model,name,sources,driver)
define model
$(OUT)/ref_$(1): $(OUT)/common.stamp $(addprefix $(REF)/,$(addsuffix .f90,$(2))) $(HERE)ref_$(1).f90
	@mkdir -p $(OUT)/$(1)
	@for f in $(2); do \
	  $(AMDFLANG) $(FFLAGS) -J$(OUT)/$(1) -I$(OUT) -c $(REF)/$$$$f.f90 -o $(OUT)/$(1)/$$$$(basename $$$$f).o || exit 1; \
	done
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/$(1) -I$(OUT) -c $(HERE)ref_$(1).f90 -o $(OUT)/$(1)/ref_$(1).o
	$(AMDFLANG) $(FFLAGS) -o $$@.tmp $(COMMON_OBJ) $(addprefix $(OUT)/$(1)/,$(addsuffix .o,$(notdir $(2)))) \
	  $(OUT)/$(1)/ref_$(1).o && mv $$@.tmp $$@
endef
$(eval $(call model,adv2d,$(ADV2D_SRC)))
$(eval $(call model,adv3d,$(ADV3D_SRC)))
$(eval $(call model,cgsem2d,$(CGSEM2D_SRC)))
$(eval $(call model,krylov,$(KRYLOV_SRC)))
$(eval $(call model,nodal,))
$(eval $(call model,pickup,$(ADV2D_SRC)))
$(eval $(call model,specmesh,))
$(eval $(call model,sw2d,$(SW2D_SRC)))

# The reference's CGSEM example driver (examples/cgsem-elliptic2d/paraboloid),
# unmodified: against the reference's own modules (ref_cgsem_paraboloid, whose
# output files are the golden files of the Fortran drop-in), and against the
# drop-in modules of self_amd/fortran linked with libself_amd.so
# (dropin_cgsem_paraboloid, run by the GPU tests; built after
# `make -C self_amd/fortran`).
CGSEM_DRIVER := ../examples/cgsem-elliptic2d/paraboloid/localmods/CGsemElliptic2D_Driver
FDROP        := $(HERE)../self_amd/fortran/_build
$(OUT)/ref_cgsem_paraboloid: $(OUT)/ref_cgsem2d $(REF)/$(CGSEM_DRIVER).f90
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/cgsem2d -I$(OUT) -c $(REF)/$(CGSEM_DRIVER).f90 -o $(OUT)/cgsem2d/paraboloid_driver.o
	$(AMDFLANG) $(FFLAGS) -o $@.tmp $(COMMON_OBJ) $(addprefix $(OUT)/cgsem2d/,$(addsuffix .o,$(notdir $(CGSEM2D_SRC)))) \
	  $(OUT)/cgsem2d/paraboloid_driver.o && mv $@.tmp $@

$(OUT)/dropin_cgsem_paraboloid: $(REF)/$(CGSEM_DRIVER).f90 $(FDROP)/cgsem_dropin.o
	@mkdir -p $(OUT)/dropin
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/dropin -I$(FDROP) -c $(REF)/$(CGSEM_DRIVER).f90 -o $(OUT)/dropin/paraboloid_driver.o
	$(AMDFLANG) $(FFLAGS) -o $@.tmp $(OUT)/dropin/paraboloid_driver.o $(FDROP)/self_amd_cabi.o \
	  $(FDROP)/shallowwater_amd.o $(FDROP)/sw_dropin.o $(FDROP)/cgsem_dropin.o \
	  -L$(HERE)../self_amd -lself_amd -Wl,-rpath,'$$ORIGIN/../../self_amd' && mv $@.tmp $@

# The 2-D advection model's own programs (src/highend/advection-diffusion2d),
# unmodified: its initial-condition program and its driver against the
# reference's modules (their files are the golden files of the Fortran
# drop-in), and the driver against the drop-in modules with libself_amd.so.
ADV2D_IC  := highend/advection-diffusion2d/Advection_InitialConditions
ADV2D_DRV := highend/advection-diffusion2d/Advection_Driver
ADV2D_OBJ := $(addprefix $(OUT)/adv2d/,$(addsuffix .o,$(notdir $(ADV2D_SRC))))
$(OUT)/ref_adv2d_ic: $(OUT)/ref_adv2d $(REF)/$(ADV2D_IC).f90
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/adv2d -I$(OUT) -c $(REF)/$(ADV2D_IC).f90 -o $(OUT)/adv2d/adv2d_ic.o
	$(AMDFLANG) $(FFLAGS) -o $@.tmp $(COMMON_OBJ) $(ADV2D_OBJ) $(OUT)/adv2d/adv2d_ic.o && mv $@.tmp $@

$(OUT)/ref_adv2d_driver: $(OUT)/ref_adv2d $(REF)/$(ADV2D_DRV).f90
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/adv2d -I$(OUT) -c $(REF)/$(ADV2D_DRV).f90 -o $(OUT)/adv2d/adv2d_driver.o
	$(AMDFLANG) $(FFLAGS) -o $@.tmp $(COMMON_OBJ) $(ADV2D_OBJ) $(OUT)/adv2d/adv2d_driver.o && mv $@.tmp $@

$(OUT)/dropin_adv2d_driver: $(REF)/$(ADV2D_DRV).f90 $(FDROP)/adv_dropin.o
	@mkdir -p $(OUT)/dropin
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/dropin -I$(FDROP) -c $(REF)/$(ADV2D_DRV).f90 -o $(OUT)/dropin/adv2d_driver.o
	$(AMDFLANG) $(FFLAGS) -o $@.tmp $(OUT)/dropin/adv2d_driver.o $(FDROP)/self_amd_cabi.o \
	  $(FDROP)/shallowwater_amd.o $(FDROP)/sw_dropin.o $(FDROP)/adv_dropin.o \
	  -L$(HERE)../self_amd -lself_amd -Wl,-rpath,'$$ORIGIN/../../self_amd' && mv $@.tmp $@

# The 3-D advection model's own programs (src/highend/advection3d), the same
# way: against the reference's modules, and against the 3-D drop-in modules
# (self_amd/fortran/_build3d, which reads the shared ones from _build).
ADV3D_OBJ := $(addprefix $(OUT)/adv3d/,$(addsuffix .o,$(notdir $(ADV3D_SRC))))
FDROP3    := $(HERE)../self_amd/fortran/_build3d
DROP3_OBJ := $(FDROP)/self_amd_cabi.o $(FDROP)/shallowwater_amd.o $(FDROP)/sw_dropin.o $(FDROP3)/adv3d_dropin.o
# $(call adv3d_prog,short name,program file under highend/advection3d)
define adv3d_prog
$(OUT)/ref_adv3d_$(1): $(OUT)/ref_adv3d $(REF)/highend/advection3d/$(2).f90
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/adv3d -I$(OUT) -c $(REF)/highend/advection3d/$(2).f90 -o $(OUT)/adv3d/prog_$(1).o
	$(AMDFLANG) $(FFLAGS) -o $$@.tmp $(COMMON_OBJ) $(ADV3D_OBJ) $(OUT)/adv3d/prog_$(1).o && mv $$@.tmp $$@

$(OUT)/dropin_adv3d_$(1): $(REF)/highend/advection3d/$(2).f90 $(FDROP3)/adv3d_dropin.o
	@mkdir -p $(OUT)/dropin3d
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/dropin3d -I$(FDROP3) -I$(FDROP) -c $(REF)/highend/advection3d/$(2).f90 \
	  -o $(OUT)/dropin3d/prog_$(1).o
	$(AMDFLANG) $(FFLAGS) -o $$@.tmp $(OUT)/dropin3d/prog_$(1).o $(DROP3_OBJ) \
	  -L$(HERE)../self_amd -lself_amd -Wl,-rpath,'$$$$ORIGIN/../../self_amd' && mv $$@.tmp $$@
endef
$(eval $(call adv3d_prog,ic,Advection3D_InitialConditions))
$(eval $(call adv3d_prog,driver,Advection3D_Driver))

# The 1-D elliptic example's driver (examples/elliptic1d/parabola/localmods/
# Elliptic1D_Driver.f90), unmodified, against the drop-in modules
# (self_amd/fortran/elliptic1d_dropin.f90) with libself_amd.so.  The
# example's own Elliptic1D_Class does not compile against the reference's
# IterativeSolvers_Class (flang: its CopyTo / CopyFrom overrides, :90-91,
# are incompatible with the deferred interfaces), so its golden files are
# the outputs the example ships under run/.
ELL1D_DIR := ../examples/elliptic1d/parabola/localmods
$(OUT)/dropin_elliptic1d_driver: $(REF)/$(ELL1D_DIR)/Elliptic1D_Driver.f90 $(FDROP)/elliptic1d_dropin.o
	@mkdir -p $(OUT)/dropin
	$(AMDFLANG) $(FFLAGS) -J$(OUT)/dropin -I$(FDROP) -c $(REF)/$(ELL1D_DIR)/Elliptic1D_Driver.f90 \
	  -o $(OUT)/dropin/elliptic1d_driver.o
	$(AMDFLANG) $(FFLAGS) -o $@.tmp $(OUT)/dropin/elliptic1d_driver.o $(FDROP)/self_amd_cabi.o \
	  $(FDROP)/shallowwater_amd.o $(FDROP)/sw_dropin.o $(FDROP)/elliptic1d_dropin.o \
	  -L$(HERE)../self_amd -lself_amd -Wl,-rpath,'$$ORIGIN/../../self_amd' && mv $@.tmp $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
