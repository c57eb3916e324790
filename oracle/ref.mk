# Builds the reference implementation (ASDMatch) from its own sources, read in place under
# $(REF), into oracle/_ref/.  Nothing here is copied into the repository: objects and binaries
# land in oracle/_ref/ only, which git ignores.  The result is test infrastructure (the golden
# fixture generator and the "reference" CPU baseline), never part of the product.
#
#   make -f oracle/ref.mk -j8            (from the repository root)
#
# Products:
#   oracle/_ref/ASDMatch.out         count build (reference defaults, config/config.h)
#   oracle/_ref/ASDMatch_local.out   local-count build: ALL_LEVEL + LOCAL_COUNT defined on the
#                                    compile line (config/config.h leaves both commented out)
#   oracle/_ref/prepare.out          executable/prepare.cpp
#   oracle/_ref/ref_driver[_local]   oracle/ref_driver.cpp: the reference's libraries driven in
#                                    main.cpp's order with the trie-root allocation moved after
#                                    reorderBags (the unmodified main.cpp indexes past the end of
#                                    its trie-root vector whenever a virtual global bag is added)
#   oracle/_ref/ref_cs_dump          oracle/ref_cs_dump.cpp linked against the reference's
#                                    relation/ + utility/ objects: prints CandidateSpace
#                                    candidate sets so the filter restatement can be pinned.
#
# GLPK 4.55 (vendored by the reference under utility/td/glpk-4.55) is compiled from its C
# sources with the configuration header the reference ships; the prebuilt objects and
# libraries that sit beside those sources are never read or linked.

REF ?= /root/reference
OUT := oracle/_ref
GLPK := $(REF)/utility/td/glpk-4.55

CC  ?= gcc
CXX ?= g++
ARCHFLAGS := -march=x86-64-v3
CFLAGS_GLPK := -O2 -w $(ARCHFLAGS) -DHAVE_CONFIG_H -I$(GLPK) \
	-I$(GLPK)/src -I$(GLPK)/src/amd -I$(GLPK)/src/bflib -I$(GLPK)/src/cglib \
	-I$(GLPK)/src/colamd -I$(GLPK)/src/env -I$(GLPK)/src/minisat -I$(GLPK)/src/misc \
	-I$(GLPK)/src/proxy -I$(GLPK)/src/zlib
CXXFLAGS_REF := -O3 -std=c++17 -w $(ARCHFLAGS) -I$(REF)/config -I$(REF)/utility -I$(REF)/relation \
	-I$(REF)/plan -I$(REF)/matching -I$(REF)/utility/td -I$(GLPK)/src

GLPK_SRCS := $(shell find $(GLPK)/src -name '*.c' 2>/dev/null | sort)
GLPK_OBJS := $(patsubst $(GLPK)/src/%.c,$(OUT)/glpk/%.o,$(GLPK_SRCS))

LIB_SRCS := utility/td/dataloader.cpp utility/td/hypergraph.cpp utility/td/preprocessing.cpp \
	utility/td/td.cpp utility/td/twbound.cpp utility/td/SCsolver.cpp \
	utility/command_parser.cpp utility/utils.cpp utility/command.cpp \
	utility/compute_set_intersection.cpp \
	relation/candidate_space.cpp relation/graph.cpp relation/relation.cpp relation/dynamic_array.cpp \
	plan/decomposition.cpp plan/estimator.cpp plan/optimizer.cpp plan/subset_structure.cpp \
	matching/join.cpp matching/adaptive.cpp

CNT_OBJS := $(patsubst %.cpp,$(OUT)/cnt/%.o,$(LIB_SRCS))
LOC_OBJS := $(patsubst %.cpp,$(OUT)/loc/%.o,$(LIB_SRCS))

all: $(OUT)/ASDMatch.out $(OUT)/ASDMatch_local.out $(OUT)/prepare.out $(OUT)/ref_cs_dump \
	$(OUT)/ref_driver $(OUT)/ref_driver_local

$(OUT)/glpk/%.o: $(GLPK)/src/%.c
	@mkdir -p $(dir $@)
	$(CC) $(CFLAGS_GLPK) -c $< -o $@

$(OUT)/libglpk_src.a: $(GLPK_OBJS)
	rm -f $@ && ar rcs $@ $^

$(OUT)/cnt/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS_REF) -c $< -o $@

$(OUT)/loc/%.o: $(REF)/%.cpp
	@mkdir -p $(dir $@)
	$(CXX) $(CXXFLAGS_REF) -DALL_LEVEL -DLOCAL_COUNT -c $< -o $@

$(OUT)/ASDMatch.out: $(REF)/main.cpp $(CNT_OBJS) $(OUT)/libglpk_src.a
	$(CXX) $(CXXFLAGS_REF) $< $(CNT_OBJS) $(OUT)/libglpk_src.a -lm -o $@

$(OUT)/ASDMatch_local.out: $(REF)/main.cpp $(LOC_OBJS) $(OUT)/libglpk_src.a
	$(CXX) $(CXXFLAGS_REF) -DALL_LEVEL -DLOCAL_COUNT $< $(LOC_OBJS) $(OUT)/libglpk_src.a -lm -o $@

$(OUT)/prepare.out: $(REF)/executable/prepare.cpp $(CNT_OBJS) $(OUT)/libglpk_src.a
	$(CXX) $(CXXFLAGS_REF) $< $(CNT_OBJS) $(OUT)/libglpk_src.a -lm -o $@

$(OUT)/ref_cs_dump: oracle/ref_cs_dump.cpp $(CNT_OBJS) $(OUT)/libglpk_src.a
	$(CXX) $(CXXFLAGS_REF) $< $(CNT_OBJS) $(OUT)/libglpk_src.a -lm -o $@

$(OUT)/ref_driver: oracle/ref_driver.cpp $(CNT_OBJS) $(OUT)/libglpk_src.a
	$(CXX) $(CXXFLAGS_REF) $< $(CNT_OBJS) $(OUT)/libglpk_src.a -lm -o $@

$(OUT)/ref_driver_local: oracle/ref_driver.cpp $(LOC_OBJS) $(OUT)/libglpk_src.a
	$(CXX) $(CXXFLAGS_REF) -DALL_LEVEL -DLOCAL_COUNT $< $(LOC_OBJS) $(OUT)/libglpk_src.a -lm -o $@

.PHONY: all
