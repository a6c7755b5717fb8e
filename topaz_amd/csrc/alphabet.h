// TOPAZ's letter tables (lexicographical.h:19-74), built once per process:
// aa_table2 (letter -> row of the 25 x 25 matrix), upper_aa and
// upper_aa_only.
#ifndef TOPAZ_AMD_ALPHABET_H
#define TOPAZ_AMD_ALPHABET_H

#include <stdint.h>
#include <string.h>

namespace tpz {

struct Alphabet {
    int8_t table2[128];       // aa_table2
    char upper[256];          // upper_aa: lower case raised, other bytes 'X'
    char upper_only[256];     // upper_aa_only: other bytes (lower case too) 'V'
    Alphabet() {
        static const char internal_order[] = "ACMLJIVTSKRQZEDBNHYFWXP*G";
        static const char matrix_order[] = "ARNDCQEGHILKMFPSTWYVBJZX*";
        for (int c = 0; c < 128; ++c) table2[c] = 23;
        for (int k = 0; k < 25; ++k) {
            const int8_t idx = (int8_t)(strchr(matrix_order, internal_order[k]) - matrix_order);
            table2['A' + k] = idx;
            table2['a' + k] = idx;
        }
        for (int c = 0; c < 256; ++c) {
            const char ch = (char)c;
            const bool keep = (ch >= 'A' && ch <= 'Y') || ch == '*' || ch == '\n';
            upper_only[c] = keep ? ch : 'V';
            if (keep) upper[c] = ch;
            else if (ch >= 'a' && ch <= 'y') upper[c] = (char)(ch - 32);
            else upper[c] = 'X';
        }
    }
    // aa_table2[(int)c]; bytes above 127 (outside the reference's table) as 127
    int8_t idx(char c) const { return table2[(int)(unsigned char)c & 127]; }
};

inline const Alphabet& alphabet() {
    static const Alphabet a;
    return a;
}

}  // namespace tpz

#endif
