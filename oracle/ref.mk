# Builds the reference's own sequential FP64 BIEF (TELEMAC-MASCARET, Fortran)
# from its sources where they lie, with AMD flang, into oracle/_ref/.
#
# This is test infrastructure only: the library it produces is the checker
# that the CPU restatement (oracle/bief_oracle.c) is validated against, and
# the "reference" CPU baseline. Nothing in the product links it.
#
# The reference is compiled without HAVE_MPI, so its P_* wrappers are its own
# void single-process stubs (sources/utils/parallel/p_dsum.F:81-87).
# No reference file is copied, generated or stubbed: the objects come from
# the .f/.F files under REFSRC as they are.
#
#   make -f oracle/ref.mk            (from the repository root)
#
REFSRC  ?= /root/reference/sources/utils
# telemac2d's friction routines (FRICTION_CALC, FRICTI) and the modules their
# interfaces need, for the PROPAG operand sequence of the driver
REFTOP  ?= $(REFSRC)/..
OUT     ?= oracle/_ref
ifeq ($(origin FC),default)
FC      := flang
endif
FFLAGS  ?= -O2 -cpp -ffixed-line-length-132 -fPIC
OBJ     := $(OUT)/obj
MOD     := $(OUT)/mod

# module files first, in dependency order
MODSRC := special/declarations_special.F \
          parallel/declarations_parallel.F parallel/interface_parallel.f \
          bief/bief_def.f bief/declarations_telemac.f bief/bief.f \
          hermes/interface_hermes.f

SPECIAL  := $(filter-out special/declarations_special.F special/gracestop.f,$(patsubst $(REFSRC)/%,%,$(wildcard $(REFSRC)/special/*.f $(REFSRC)/special/*.F)))
PARALLEL := $(filter-out parallel/declarations_parallel.F parallel/interface_parallel.f,$(patsubst $(REFSRC)/%,%,$(wildcard $(REFSRC)/parallel/*.f $(REFSRC)/parallel/*.F)))
HERMES   := $(filter-out hermes/interface_hermes.f,$(patsubst $(REFSRC)/%,%,$(wildcard $(REFSRC)/hermes/*.f $(REFSRC)/hermes/*.F)))
# mt02pt.f is off the hot path and rejected by flang (rank-mismatch PARAMETER)
BIEFALL  := $(filter-out bief/bief_def.f bief/declarations_telemac.f bief/bief.f bief/mt02pt.f,$(patsubst $(REFSRC)/%,%,$(wildcard $(REFSRC)/bief/*.f $(REFSRC)/bief/*.F)))

# modules that some BIEF routines use and that live in ordinary-looking files
BIEFMODS := bief/streamline.f bief/algae_transp.f
BIEFREST := $(filter-out $(BIEFMODS),$(BIEFALL))

objname = $(OBJ)/$(subst /,__,$(basename $(1))).o

MODOBJ  := $(foreach s,$(MODSRC),$(call objname,$(s)))
BMODOBJ := $(foreach s,$(BIEFMODS),$(call objname,$(s)))
# hermes (mesh/result file I/O) is off the path: only its interface module is
# compiled, because some BIEF I/O routines name it
RESTOBJ := $(foreach s,$(SPECIAL) $(PARALLEL) $(BIEFREST),$(call objname,$(s)))

all: $(OUT)/libbief_ref.so

$(OUT)/stamp_mods: | $(OBJ) $(MOD)
	@set -e; for s in $(MODSRC); do \
	  o=$(OBJ)/$$(echo $${s%.*} | sed 's#/#__#g').o; \
	  $(FC) $(FFLAGS) -module-dir $(MOD) -I$(MOD) -c $(REFSRC)/$$s -o $$o; \
	done; \
	for s in $(BIEFMODS); do \
	  o=$(OBJ)/$$(echo $${s%.*} | sed 's#/#__#g').o; \
	  $(FC) $(FFLAGS) -module-dir $(MOD) -I$(MOD) -c $(REFSRC)/$$s -o $$o; \
	done; touch $@

$(OBJ)/%.o: $(OUT)/stamp_mods
	@s=$$(echo $* | sed 's#__#/#'); f=$$(ls $(REFSRC)/$$s.f $(REFSRC)/$$s.F 2>/dev/null | head -1); \
	$(FC) $(FFLAGS) -module-dir $(OBJ)/m_$* -I$(MOD) -c $$f -o $@ 2>$(OBJ)/$*.err || (echo "skip $$s"; rm -f $@); \
	rm -rf $(OBJ)/m_$*

# module files in dependency order, then the routines (DRAGFO: FRICTI names it)
T2DMODS := waqtel/declarations_waqtel.f telemac2d/friction_def.f \
           telemac2d/declarations_telemac2d.f telemac2d/interface_telemac2d.f
T2DSRCS := telemac2d/friction_calc.f telemac2d/fricti.f telemac2d/dragfo.f

$(OUT)/stamp_t2d: $(OUT)/stamp_mods
	@set -e; for s in $(T2DMODS) $(T2DSRCS); do \
	  o=$(OBJ)/t2d__$$(basename $${s%.*}).o; \
	  $(FC) $(FFLAGS) -module-dir $(MOD) -I$(MOD) -c $(REFTOP)/$$s -o $$o; \
	done; touch $@

$(OUT)/libbief_ref.a: $(OUT)/stamp_mods $(OUT)/stamp_t2d $(RESTOBJ)
	@rm -f $@; ar rcs $@ $(sort $(MODOBJ) $(BMODOBJ) $(wildcard $(OBJ)/*.o))

$(OBJ) $(MOD):
	mkdir -p $@

.PHONY: all

# project-written driver (bind(C) entry points) linked against the archive
$(OUT)/libbief_ref.so: oracle/ref_driver.f90 $(OUT)/libbief_ref.a
	$(FC) $(FFLAGS) -module-dir $(OUT) -I$(MOD) -shared oracle/ref_driver.f90 $(OUT)/libbief_ref.a -o $@

so: $(OUT)/libbief_ref.so
.PHONY: so
