/*
 * abi_layout.cpp - this library's view of the PG-Strom ABI structs, dumped
 * as JSON with the same keys as the reference dump (tests/golden/
 * abi_layout.json) so the conformance test can compare them one by one.
 */
#include <stdio.h>
#include <string>

#include "xpuhip.h"
#include "xpu_opcodes_table.h"

namespace {
std::string g_json;

void emit(const char *key, long long v)
{
    char buf[256];
    snprintf(buf, sizeof(buf), "%s\"%s\": %lld", g_json.size() > 1 ? ",\n " : "", key, v);
    g_json += buf;
}
}   /* namespace */

#define SZ(T)               emit("sizeof:" #T, (long long)sizeof(T))
#define OFF(T, F)           emit("offsetof:" #T "." #F, (long long)offsetof(T, F))
#define OFFK(KEY, T, F)     emit("offsetof:" KEY, (long long)offsetof(T, F))

extern "C" const char *
xpuhip_abi_layout_json(void)
{
    g_json = "{";
    SZ(kern_errorbuf);
    OFF(kern_errorbuf, lineno); OFF(kern_errorbuf, filename);
    OFF(kern_errorbuf, funcname); OFF(kern_errorbuf, message);
    SZ(ArrowTypeOptions);
    OFF(ArrowTypeOptions, unitsz); OFF(ArrowTypeOptions, align);
    OFFK("ArrowTypeOptions.integer.bitWidth", ArrowTypeOptions, u.integer.bitWidth);
    OFFK("ArrowTypeOptions.integer.is_signed", ArrowTypeOptions, u.integer.is_signed);
    OFFK("ArrowTypeOptions.floating_point.precision", ArrowTypeOptions, u.floating_point.precision);
    SZ(kern_colmeta);
    OFF(kern_colmeta, attbyval); OFF(kern_colmeta, attalign); OFF(kern_colmeta, attlen);
    OFF(kern_colmeta, attnum); OFF(kern_colmeta, attcacheoff); OFF(kern_colmeta, atttypid);
    OFF(kern_colmeta, atttypmod); OFF(kern_colmeta, atttypkind); OFF(kern_colmeta, kds_format);
    OFF(kern_colmeta, dtype_sizeof); OFF(kern_colmeta, kds_offset);
    OFF(kern_colmeta, idx_subattrs); OFF(kern_colmeta, num_subattrs);
    OFF(kern_colmeta, attname); OFF(kern_colmeta, attopts); OFF(kern_colmeta, field_index);
    OFF(kern_colmeta, virtual_offset); OFF(kern_colmeta, nullmap_offset);
    OFF(kern_colmeta, nullmap_length); OFF(kern_colmeta, values_offset);
    OFF(kern_colmeta, values_length); OFF(kern_colmeta, extra_offset);
    OFF(kern_colmeta, extra_length);
    SZ(kern_data_store);
    OFF(kern_data_store, length); OFF(kern_data_store, usage); OFF(kern_data_store, nitems);
    OFF(kern_data_store, ncols); OFF(kern_data_store, format); OFF(kern_data_store, has_varlena);
    OFF(kern_data_store, tdhasoid); OFF(kern_data_store, tdtypeid); OFF(kern_data_store, tdtypmod);
    OFF(kern_data_store, table_oid); OFF(kern_data_store, hash_nslots);
    OFF(kern_data_store, block_offset); OFF(kern_data_store, block_nloaded);
    OFF(kern_data_store, column_nrooms); OFF(kern_data_store, arrow_virtual_usage);
    OFF(kern_data_store, parquet_row_group); OFF(kern_data_store, parquet_parallel_load);
    OFF(kern_data_store, parquet_cache_enabled); OFF(kern_data_store, nr_colmeta);
    OFF(kern_data_store, colmeta);
    SZ(kern_tupitem);
    OFF(kern_tupitem, t_len); OFF(kern_tupitem, hash); OFF(kern_tupitem, has_xact_attrs);
    OFF(kern_tupitem, t_infomask2); OFF(kern_tupitem, t_infomask); OFF(kern_tupitem, t_hoff);
    OFF(kern_tupitem, t_bits);
    emit("const:MINIMAL_TUPLE_OFFSET", MINIMAL_TUPLE_OFFSET);
    SZ(kern_hashitem); OFF(kern_hashitem, next); OFF(kern_hashitem, t);
    SZ(kern_fallbackitem); OFF(kern_fallbackitem, depth); OFF(kern_fallbackitem, matched);
    OFF(kern_fallbackitem, t);
    SZ(kern_varload_desc); OFF(kern_varload_desc, vl_slot_id);
    SZ(kern_varmove_desc); OFF(kern_varmove_desc, vm_slot_id); OFF(kern_varmove_desc, vm_from_xdatum);
    SZ(kern_varslot_desc);
    OFF(kern_varslot_desc, vs_typbyval); OFF(kern_varslot_desc, vs_typalign);
    OFF(kern_varslot_desc, vs_typlen); OFF(kern_varslot_desc, vs_typmod);
    OFF(kern_varslot_desc, vs_offset); OFF(kern_varslot_desc, idx_subfield);
    OFF(kern_varslot_desc, num_subfield); OFF(kern_varslot_desc, vs_ops);
    SZ(kern_fallback_desc); OFF(kern_fallback_desc, fb_slot_id); OFF(kern_fallback_desc, fb_kvec_offset);
    SZ(kern_aggregate_desc);
    OFF(kern_aggregate_desc, arg0_slot_id); OFF(kern_aggregate_desc, arg1_slot_id);
    OFF(kern_aggregate_desc, filter_slot_id); OFF(kern_aggregate_desc, typmod);
    OFF(kern_expression, len); OFF(kern_expression, exptype); OFF(kern_expression, expr_ops);
    OFF(kern_expression, expflags); OFF(kern_expression, opcode); OFF(kern_expression, fn_dptr);
    OFF(kern_expression, nr_args); OFF(kern_expression, args_offset); OFF(kern_expression, u.data);
    OFF(kern_expression, u.c.const_type); OFF(kern_expression, u.c.const_isnull);
    OFF(kern_expression, u.c.const_value);
    OFF(kern_expression, u.p.param_id); OFF(kern_expression, u.p.__data);
    OFF(kern_expression, u.v.var_offset); OFF(kern_expression, u.v.var_slot_id);
    OFF(kern_expression, u.v.__data);
    OFF(kern_expression, u.casewhen.case_comp); OFF(kern_expression, u.casewhen.case_else);
    OFF(kern_expression, u.casewhen.data);
    OFF(kern_expression, u.saop.elem_slot_id); OFF(kern_expression, u.saop.data);
    OFF(kern_expression, u.haop.elem_ops); OFF(kern_expression, u.haop.elem_type);
    OFF(kern_expression, u.haop.nslots); OFF(kern_expression, u.haop.default_offset);
    OFF(kern_expression, u.haop.hslots);
    SZ(kern_hashed_array_op_item); OFF(kern_hashed_array_op_item, next);
    OFF(kern_hashed_array_op_item, data);
    emit("const:KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY", KERN_HASHED_ARRAY_OP_ITEM_HAS_KEY);
    emit("const:KERN_HASHED_ARRAY_OP_ITEM_HAS_VALUE", KERN_HASHED_ARRAY_OP_ITEM_HAS_VALUE);
    emit("const:KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK", KERN_HASHED_ARRAY_OP_ITEM_NEXT_MASK);
    emit("const:KERN_HASHED_ARRAY_OP_NITEMS_THRESHOLD", KERN_HASHED_ARRAY_OP_NITEMS_THRESHOLD);
    OFF(kern_expression, u.load.depth); OFF(kern_expression, u.load.nitems);
    OFF(kern_expression, u.load.desc);
    OFF(kern_expression, u.move.depth); OFF(kern_expression, u.move.nitems);
    OFF(kern_expression, u.move.desc);
    OFF(kern_expression, u.save.sv_slot_id); OFF(kern_expression, u.save.data);
    OFF(kern_expression, u.pagg.nattrs); OFF(kern_expression, u.pagg.desc);
    OFF(kern_expression, u.proj.hash); OFF(kern_expression, u.proj.nattrs);
    OFF(kern_expression, u.proj.slot_id);
    OFF(kern_expression, u.pack.npacked); OFF(kern_expression, u.pack.offset);
    SZ(kagg_state__pminmax_int64_packed); OFF(kagg_state__pminmax_int64_packed, value);
    SZ(kagg_state__pminmax_fp64_packed); OFF(kagg_state__pminmax_fp64_packed, value);
    SZ(kagg_state__psum_int_packed); OFF(kagg_state__psum_int_packed, sum);
    SZ(kagg_state__psum_fp_packed); OFF(kagg_state__psum_fp_packed, sum);
    SZ(kagg_state__psum_numeric_packed); OFF(kagg_state__psum_numeric_packed, sum);
    SZ(kagg_state__stddev_packed); SZ(kagg_state__covar_packed);
    SZ(kern_session_info);
    OFF(kern_session_info, query_plan_id); OFF(kern_session_info, kcxt_kvars_nrooms);
    OFF(kern_session_info, kcxt_kvars_nslots); OFF(kern_session_info, kcxt_kvars_defs);
    OFF(kern_session_info, kcxt_kvecs_bufsz); OFF(kern_session_info, kcxt_kvecs_ndims);
    OFF(kern_session_info, kcxt_extra_bufsz); OFF(kern_session_info, cuda_stack_size);
    OFF(kern_session_info, xpu_task_flags); OFF(kern_session_info, optimal_gpus);
    OFF(kern_session_info, xpucode_load_vars_packed); OFF(kern_session_info, xpucode_move_vars_packed);
    OFF(kern_session_info, xpucode_scan_quals); OFF(kern_session_info, xpucode_join_quals_packed);
    OFF(kern_session_info, xpucode_hash_values_packed); OFF(kern_session_info, xpucode_gist_evals_packed);
    OFF(kern_session_info, xpucode_projection); OFF(kern_session_info, xpucode_groupby_keyhash);
    OFF(kern_session_info, xpucode_groupby_keyload); OFF(kern_session_info, xpucode_groupby_keycomp);
    OFF(kern_session_info, xpucode_groupby_actions); OFF(kern_session_info, xpucode_gpusort_keydesc);
    OFF(kern_session_info, hostEpochTimestamp); OFF(kern_session_info, session_xact_state);
    OFF(kern_session_info, session_timezone); OFF(kern_session_info, session_encode);
    OFF(kern_session_info, session_currency_frac_digits);
    OFF(kern_session_info, pinned_inner_buffer_partition_size);
    OFF(kern_session_info, projection_kds_dst); OFF(kern_session_info, select_into_pathname);
    OFF(kern_session_info, select_into_kds_head); OFF(kern_session_info, select_into_projdesc);
    OFF(kern_session_info, pgsql_port_number); OFF(kern_session_info, pgsql_plan_node_id);
    OFF(kern_session_info, join_inner_handle); OFF(kern_session_info, groupby_kds_final);
    OFF(kern_session_info, groupby_prepfn_bufsz); OFF(kern_session_info, groupby_ngroups_estimation);
    OFF(kern_session_info, gpusort_htup_margin); OFF(kern_session_info, gpusort_limit_count);
    OFF(kern_session_info, fallback_kds_head); OFF(kern_session_info, fallback_desc_defs);
    OFF(kern_session_info, fallback_desc_nitems); OFF(kern_session_info, nparams);
    OFF(kern_session_info, poffset);
    SZ(kern_exec_task); OFF(kern_exec_task, kds_src_pathname); OFF(kern_exec_task, kds_src_iovec);
    OFF(kern_exec_task, kds_src_offset); OFF(kern_exec_task, data);
    SZ(kern_exec_results);
    OFF(kern_exec_results, chunks_offset); OFF(kern_exec_results, chunks_nitems);
    OFF(kern_exec_results, ojmap_offset); OFF(kern_exec_results, ojmap_length);
    OFF(kern_exec_results, right_outer_join); OFF(kern_exec_results, final_plan_task);
    OFF(kern_exec_results, final_nitems); OFF(kern_exec_results, final_usage);
    OFF(kern_exec_results, final_total); OFF(kern_exec_results, final_sorting_msec);
    OFF(kern_exec_results, select_into_nblocks); OFF(kern_exec_results, npages_direct_read);
    OFF(kern_exec_results, nitems_raw); OFF(kern_exec_results, nitems_in);
    OFF(kern_exec_results, nitems_out); OFF(kern_exec_results, num_rels);
    OFF(kern_exec_results, stats);
    OFF(XpuCommand, magic); OFF(XpuCommand, tag); OFF(XpuCommand, length);
    OFF(XpuCommand, priv); OFF(XpuCommand, chain); OFF(XpuCommand, u);
    SZ(kern_multirels);
    OFF(kern_multirels, length); OFF(kern_multirels, ojmap_sz); OFF(kern_multirels, kbuf_part_offset);
    OFF(kern_multirels, num_inner_rels); OFF(kern_multirels, chunks);
    OFF(kern_multirels, chunks[0].kds_offset); OFF(kern_multirels, chunks[0].ojmap_offset);
    OFF(kern_multirels, chunks[0].gist_offset); OFF(kern_multirels, chunks[0].gist_ctid_resno);
    OFF(kern_multirels, chunks[0].is_nestloop); OFF(kern_multirels, chunks[0].left_outer);
    OFF(kern_multirels, chunks[0].right_outer); OFF(kern_multirels, chunks[0].pinned_buffer);
    OFF(kern_multirels, chunks[0].buffer_id);
    emit("sizeof:kern_multirels.chunks[0]", (long long)sizeof(kern_multirels_chunk));
    /* opcode numbers this build executes */
#define XT(NAME)   emit("type:" #NAME, TypeOpCode__##NAME)
    XT(bool); XT(int1); XT(int2); XT(int4); XT(int8); XT(float2); XT(float4); XT(float8);
    XT(numeric); XT(bytea); XT(text); XT(bpchar); XT(date);
#define XF(NAME)   emit("func:" #NAME, FuncOpCode__##NAME)
    XF(ConstExpr); XF(ParamExpr); XF(VarExpr); XF(BoolExpr_And); XF(BoolExpr_Or);
    XF(BoolExpr_Not); XF(NullTestExpr_IsNull); XF(NullTestExpr_IsNotNull);
    XF(BoolTestExpr_IsTrue); XF(BoolTestExpr_IsNotTrue); XF(BoolTestExpr_IsFalse);
    XF(BoolTestExpr_IsNotFalse); XF(BoolTestExpr_IsUnknown); XF(BoolTestExpr_IsNotUnknown);
    XF(DistinctFrom); XF(CoalesceExpr); XF(LeastExpr); XF(GreatestExpr); XF(CaseWhenExpr);
    XF(LoadVars); XF(MoveVars); XF(JoinQuals); XF(HashValue); XF(GiSTEval); XF(SaveExpr);
    XF(AggFuncs); XF(Projection); XF(SortKeys); XF(Packed);
    XF(ScalarArrayOpAny); XF(ScalarArrayOpAll); XF(HashedArrayOp);
    XT(array);
#define XFUNC_EMIT(OPC, NAME, KIND, OP, NARGS, A0, A1, A2, RES) emit("func:" #NAME, OPC);
    XPU_FUNCTION_TABLE(XFUNC_EMIT)
    emit("const:KERN_EXPRESSION_MAGIC", KERN_EXPRESSION_MAGIC);
    g_json += "}\n";
    return g_json.c_str();
}
