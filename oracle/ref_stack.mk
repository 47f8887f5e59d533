# The reference's Core::StackAllocator, compiled from its own source where it
# lies, with the C shim of this folder: oracle/_ref/libref_stack.so (test
# infrastructure, the checker of oracle/stack_alloc.py).
#   make -C oracle -f ref_stack.mk REF=/root/reference
# -include cstddef: the reference's Allocator.hpp uses std::max_align_t and
# std::byte without including <cstddef> (MSVC's headers bring it in).
REF ?= /root/reference
CXX ?= g++
CXXFLAGS := -O2 -fPIC -std=c++20 -include cstddef -I$(REF)/src -Wall -Wno-delete-non-virtual-dtor
OUT := _ref/libref_stack.so
SRC := $(REF)/src/Core/Allocator/StackAllocator.cpp

all: $(OUT)

$(OUT): ref_stack_shim.cpp $(SRC)
	mkdir -p _ref
	$(CXX) $(CXXFLAGS) -shared -o $@.tmp ref_stack_shim.cpp $(SRC)
	mv $@.tmp $@

.PHONY: all
