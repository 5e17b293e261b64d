/* The NAND/STO program model shared by the host library and the GPU engine.

   A program is a straight list of 1-bit instructions over 1-bit cells:
     NAND a   RR <- !(RR & M[a])
     STO  a   M[a] <- RR            (ignored for read-only cells)
   Cell 0 reads 0 and cell 1 reads 1. Lane-id cells IDd read bit d of the machine number and link
   cells read another cell of the machine across one cube dimension (machine k XOR 2^d).

   The structs below are layout-identical to the reference's Instr / CellInfo / Program
   (src/isa.h:14-44), so a caller holding a reference Program can pass it by pointer. */
#ifndef TESSERACT_ISA_H
#define TESSERACT_ISA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { TS_OP_NAND = 0, TS_OP_STO = 1 };
enum { TS_CELL_ZERO = 0, TS_CELL_ONE = 1, TS_CELL_FIRST_FREE = 2 };

/* A group is 512 machines (one 64-byte line per cell); a cube has up to 2^16 machines. */
enum { TS_CUBE_DIMS = 9, TS_CUBE_SIZE = 1 << TS_CUBE_DIMS, TS_CUBE_MAX_DIMS = 16,
       TS_CUBE_MAX_GROUPS = 1 << (TS_CUBE_MAX_DIMS - TS_CUBE_DIMS) };

enum { TS_CK_PLAIN = 0, TS_CK_ID = 1, TS_CK_LINK = 2 };

typedef struct {
    uint8_t  op;    /* TS_OP_NAND or TS_OP_STO */
    uint32_t addr;  /* cell address */
} ts_instr;

typedef struct {
    uint8_t  kind;  /* TS_CK_* */
    uint8_t  dim;   /* id bit, or the dimension a link crosses */
    uint32_t src;   /* link: the cell read in the neighbour */
} ts_cellinfo;

typedef struct {
    ts_instr    *code;
    size_t       len;
    uint32_t     ncells;      /* including ZERO and ONE */
    char       **cell_names;  /* cell_names[addr] */
    uint32_t    *inputs;      /* .in cells in declaration order */
    size_t       ninputs;
    uint32_t    *outputs;     /* .out cells in declaration order */
    size_t       noutputs;
    ts_cellinfo *special;     /* per cell; NULL when the program has no id or link cells */
} ts_program;

/* One 64-byte line: 512 machines, machine k is bit k % 64 of w[k / 64] (reference src/lanes.h:13). */
typedef struct {
    uint64_t w[8];
} ts_line;

#ifdef __cplusplus
}
#endif

#endif
