# Builds the reference's single-threaded C++ CPU engine (`gagp`) straight from
# its sources with g++ -- no CMake, no CUDA, no OpenSSL (the grammar identity
# hash then uses the reference's own built-in SHA-256).  Outputs go only to
# oracle/_ref/.  Used to write the golden fixtures of the parity tests, and as
# the CPU baseline ("kind": "reference") timed by bench.py.
#
#   make -f oracle/ref.mk -j8            (from the repository root)
REF ?= /root/reference/cpp
OUT := oracle/_ref
SRCS := $(filter-out $(REF)/src/runtime/gpu/%,$(shell find $(REF)/src -name '*.cpp' 2>/dev/null | sort))
OBJS := $(patsubst $(REF)/src/%.cpp,$(OUT)/obj/%.o,$(SRCS))
CXXFLAGS ?= -std=c++17 -O3 -DNDEBUG -ffp-contract=off
# ref_fitness: the fitness-reduction known-answer driver (oracle/ref_fitness.cpp,
# test infrastructure) over the same objects, without the command line's main.
LIB_OBJS := $(filter-out $(OUT)/obj/cli/gagp_main.o,$(OBJS))
all: $(OUT)/gagp $(OUT)/ref_fitness
$(OUT)/obj/%.o: $(REF)/src/%.cpp
	@mkdir -p $(dir $@)
	g++ $(CXXFLAGS) -I$(REF)/include -c $< -o $@
$(OUT)/gagp: $(OBJS)
	g++ $(CXXFLAGS) $(OBJS) -o $@ -lpthread
$(OUT)/ref_fitness: oracle/ref_fitness.cpp $(LIB_OBJS)
	g++ $(CXXFLAGS) -I$(REF)/include oracle/ref_fitness.cpp $(LIB_OBJS) -o $@ -lpthread
clean:
	rm -rf $(OUT)
.PHONY: all clean
