// Embeds the Joe--Kuo 6.21201 direction asset generated by
// pricing_amd/sobol_directions.py (the reference embeds the same bytes with
// include_bytes!, crates/pricing/src/engine/sampling/qmc.rs:10).
    .section .rodata
    .balign 16
    .globl pe_jk_directions_begin
    .hidden pe_jk_directions_begin
pe_jk_directions_begin:
    .incbin PE_DIRECTIONS_FILE
    .globl pe_jk_directions_end
    .hidden pe_jk_directions_end
pe_jk_directions_end:
    .section .note.GNU-stack,"",@progbits
