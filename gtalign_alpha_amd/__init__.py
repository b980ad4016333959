"""MI355X-native structure-alignment stage of GTalign (query chunk x reference chunk ->
TM-scores, RMSD, superposition and alignment), behind the C ABI of include/gta/gtalign_amd.h."""
