# Builds the reference QMcBeaver evaluate path from its own sources, where they
# lie under /root/reference, into oracle/_ref/ (git-ignored).  Only g++ on the
# reference's .cpp files plus this project's driver (ref_driver.cpp); the
# reference's own configure.py/Makefile are not used.  The in-tree GEMM
# (Array2D.h:425-504) and Numerical-Recipes LU (Array2D.h:1521-1548) are the
# FP64 path compiled here (no -DUSEBLAS / -DUSELAPACK / -DPARALLEL).
#
#   make -f oracle/ref.mk -j8          (from the repo root)
REFROOT ?= /root/reference/QMcBeaver
OUT     := oracle/_ref
OBJDIR  := $(OUT)/obj
CXX     := g++
REFFLAGS := -std=gnu++98 -O2 -w -fpermissive -DVERSION=20090214 \
  -include cstdio -include cstring -include cstdlib -include climits \
  -include ctime -include cmath -include unistd.h -include sys/time.h \
  -include algorithm -I$(REFROOT)/include

SRCS := $(filter-out $(REFROOT)/src/QMcBeaver.cpp,$(wildcard $(REFROOT)/src/*.cpp))
OBJS := $(patsubst $(REFROOT)/src/%.cpp,$(OBJDIR)/%.o,$(SRCS))

all: $(OUT)/qmcb_ref_driver

# the reference program itself (its main, QMcBeaver.cpp, over the same
# objects): whole VMC/DMC runs of a .ckmf input, as the reference runs them
qmcbeaver: $(OUT)/QMcBeaver.x

$(OBJDIR)/%.o: $(REFROOT)/src/%.cpp
	@mkdir -p $(OBJDIR)
	$(CXX) $(REFFLAGS) -c $< -o $@

$(OUT)/qmcb_ref_driver: oracle/ref_driver.cpp $(OBJS)
	$(CXX) $(REFFLAGS) oracle/ref_driver.cpp $(OBJS) -o $@

$(OUT)/QMcBeaver.x: $(REFROOT)/src/QMcBeaver.cpp $(OBJS)
	$(CXX) $(REFFLAGS) $(REFROOT)/src/QMcBeaver.cpp $(OBJS) -o $@

.PHONY: all qmcbeaver
