/*
 * ref_driver.cc - test-only checker built against the reference sources.
 *
 * This file is compiled, together with the reference's device-agnostic VM
 * sources (src/xpu_common.cu, xpu_basetype.cu, xpu_numeric.cu, xpu_textlib.cu,
 * xpu_timelib.cu, xpu_misclib.cu, xpu_jsonlib.cu, compiled as host C++), into
 * oracle/_ref/libpgstrom_ref.so by oracle/Makefile.  It is never part of the
 * product: only tests and tests/golden generators load it, to
 *   - dump the reference's struct layouts and opcode numbers (ref_abi_json),
 *   - compute the reference's pg_hash_any (xpu_common.cu:2920-3041),
 *   - deform heap tuples with the reference's ExecLoadVarsOuterHeap.
 */
#include "xpu_common.h"
#include <stdlib.h>
#include <string>
#include <vector>

static std::string g_json;

static void emit(const char *key, long long val)
{
    char buf[256];
    snprintf(buf, sizeof(buf), "%s\"%s\": %lld", g_json.size() > 1 ? ",\n " : "", key, val);
    g_json += buf;
}

#define SZ(T)           emit("sizeof:" #T, (long long)sizeof(T))
#define OFF(T, F)       emit("offsetof:" #T "." #F, (long long)offsetof(T, F))

extern "C" const char *
ref_abi_json(void)
{
    g_json = "{";
    SZ(kern_errorbuf);
    OFF(kern_errorbuf, lineno); OFF(kern_errorbuf, filename);
    OFF(kern_errorbuf, funcname); OFF(kern_errorbuf, message);

    SZ(ArrowTypeOptions);
    OFF(ArrowTypeOptions, unitsz); OFF(ArrowTypeOptions, align);
    OFF(ArrowTypeOptions, integer.bitWidth); OFF(ArrowTypeOptions, integer.is_signed);
    OFF(ArrowTypeOptions, floating_point.precision);

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
    emit("const:MINIMAL_TUPLE_OFFSET", (long long)MINIMAL_TUPLE_OFFSET);
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
    OFF(kern_expression, u.load.depth); OFF(kern_expression, u.load.nitems);
    OFF(kern_expression, u.load.desc);
    OFF(kern_expression, u.move.depth); OFF(kern_expression, u.move.nitems);
    OFF(kern_expression, u.move.desc);
    OFF(kern_expression, u.save.sv_slot_id); OFF(kern_expression, u.save.data);
    OFF(kern_expression, u.pagg.nattrs); OFF(kern_expression, u.pagg.desc);
    OFF(kern_expression, u.proj.hash); OFF(kern_expression, u.proj.nattrs);
    OFF(kern_expression, u.proj.slot_id);
    OFF(kern_expression, u.pack.npacked); OFF(kern_expression, u.pack.offset);
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
    emit("type:composite", TypeOpCode__composite);
    emit("type:array", TypeOpCode__array);
    emit("type:internal", TypeOpCode__internal);

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
    emit("sizeof:kern_multirels.chunks[0]", (long long)sizeof(((kern_multirels *)0)->chunks[0]));

    /* opcode numbers */
#define TYPE_OPCODE(NAME,a,b)   emit("type:" #NAME, (long long)TypeOpCode__##NAME);
#include "xpu_opcodes.h"
#undef TYPE_OPCODE
#define FUNC_OPCODE(a,b,c,NAME,d,e)         emit("func:" #NAME, (long long)FuncOpCode__##NAME);
#define DEVONLY_FUNC_OPCODE(a,NAME,b,c,d)   emit("func:" #NAME, (long long)FuncOpCode__##NAME);
#include "xpu_opcodes.h"
#undef FUNC_OPCODE
#undef DEVONLY_FUNC_OPCODE
    emit("func:ConstExpr", FuncOpCode__ConstExpr);
    emit("func:ParamExpr", FuncOpCode__ParamExpr);
    emit("func:VarExpr", FuncOpCode__VarExpr);
    emit("func:BoolExpr_And", FuncOpCode__BoolExpr_And);
    emit("func:BoolExpr_Or", FuncOpCode__BoolExpr_Or);
    emit("func:BoolExpr_Not", FuncOpCode__BoolExpr_Not);
    emit("func:NullTestExpr_IsNull", FuncOpCode__NullTestExpr_IsNull);
    emit("func:NullTestExpr_IsNotNull", FuncOpCode__NullTestExpr_IsNotNull);
    emit("func:BoolTestExpr_IsTrue", FuncOpCode__BoolTestExpr_IsTrue);
    emit("func:BoolTestExpr_IsNotTrue", FuncOpCode__BoolTestExpr_IsNotTrue);
    emit("func:BoolTestExpr_IsFalse", FuncOpCode__BoolTestExpr_IsFalse);
    emit("func:BoolTestExpr_IsNotFalse", FuncOpCode__BoolTestExpr_IsNotFalse);
    emit("func:BoolTestExpr_IsUnknown", FuncOpCode__BoolTestExpr_IsUnknown);
    emit("func:BoolTestExpr_IsNotUnknown", FuncOpCode__BoolTestExpr_IsNotUnknown);
    emit("func:DistinctFrom", FuncOpCode__DistinctFrom);
    emit("func:CoalesceExpr", FuncOpCode__CoalesceExpr);
    emit("func:LeastExpr", FuncOpCode__LeastExpr);
    emit("func:GreatestExpr", FuncOpCode__GreatestExpr);
    emit("func:CaseWhenExpr", FuncOpCode__CaseWhenExpr);
    emit("func:ScalarArrayOpAny", FuncOpCode__ScalarArrayOpAny);
    emit("func:ScalarArrayOpAll", FuncOpCode__ScalarArrayOpAll);
    emit("func:HashedArrayOp", FuncOpCode__HashedArrayOp);
    emit("func:LoadVars", FuncOpCode__LoadVars);
    emit("func:MoveVars", FuncOpCode__MoveVars);
    emit("func:JoinQuals", FuncOpCode__JoinQuals);
    emit("func:HashValue", FuncOpCode__HashValue);
    emit("func:GiSTEval", FuncOpCode__GiSTEval);
    emit("func:SaveExpr", FuncOpCode__SaveExpr);
    emit("func:AggFuncs", FuncOpCode__AggFuncs);
    emit("func:Projection", FuncOpCode__Projection);
    emit("func:SortKeys", FuncOpCode__SortKeys);
    emit("func:Packed", FuncOpCode__Packed);
    emit("const:KERN_EXPRESSION_MAGIC", KERN_EXPRESSION_MAGIC);
    g_json += "}\n";
    return g_json.c_str();
}

/* The reference's own hash of a byte string (xpu_common.cu:2920-3041). */
extern "C" uint32_t
ref_pg_hash_any(const void *ptr, int sz)
{
    return pg_hash_any(ptr, sz);
}

/*
 * The reference's own heap-tuple deform: ExecLoadVarsOuterHeap
 * (xpu_common.cu:1587-1620 -> kern_extract_heap_tuple :194-257 ->
 * __kern_extract_common_tuple :60-130) on the tuple at htup of a kds whose
 * colmeta describes the columns.  Loads nload attributes resnos[i] of type
 * codes types[i] (bool, int1/2/4/8, float4/8, text); for each writes
 * nulls[i], and vals[i] (the value's bits, sign-extended integers) or, for
 * text, offs[i]/lens[i] (offset of the data from htup and its length; -1
 * for a compressed or external datum).  Returns 0, or -1 if the reference
 * raised an error.
 */
extern "C" int
ref_heap_deform(const void *kds, const void *htup, int nload, const int16_t *resnos,
                const int32_t *types, int64_t *vals, int64_t *offs, int32_t *lens,
                uint8_t *nulls)
{
    if (nload <= 0 || nload > 64)
        return -1;
    std::vector<kern_varslot_desc> desc(nload);
    std::vector<xpu_datum_t *> slots(nload);
    std::vector<std::vector<char>> store(nload);
    for (int i = 0; i < nload; i++)
    {
        const xpu_datum_operators *ops;
        switch (types[i])
        {
            case TypeOpCode__bool: ops = &xpu_bool_ops; break;
            case TypeOpCode__int1: ops = &xpu_int1_ops; break;
            case TypeOpCode__int2: ops = &xpu_int2_ops; break;
            case TypeOpCode__int4: ops = &xpu_int4_ops; break;
            case TypeOpCode__int8: ops = &xpu_int8_ops; break;
            case TypeOpCode__float4: ops = &xpu_float4_ops; break;
            case TypeOpCode__float8: ops = &xpu_float8_ops; break;
            case TypeOpCode__text: ops = &xpu_text_ops; break;
            default: return -1;
        }
        memset(&desc[i], 0, sizeof(kern_varslot_desc));
        desc[i].vs_type_code = ops->xpu_type_code;
        desc[i].vs_typbyval = ops->xpu_type_byval;
        desc[i].vs_typalign = ops->xpu_type_align;
        desc[i].vs_typlen = ops->xpu_type_length;
        desc[i].vs_offset = -1;
        desc[i].vs_ops = ops;
        store[i].assign(ops->xpu_type_sizeof + 16, 0);
        slots[i] = (xpu_datum_t *)store[i].data();
        slots[i]->expr_ops = NULL;
    }
    std::vector<char> kbuf(offsetof(kern_context, vlbuf) + 1024, 0);
    kern_context *kcxt = (kern_context *)kbuf.data();
    kcxt->kvars_slot = slots.data();
    kcxt->kvars_desc = desc.data();
    kcxt->kvars_nslots = nload;
    kcxt->kvars_nrooms = nload;
    kcxt->vlpos = kcxt->vlbuf;
    kcxt->vlend = kcxt->vlbuf + 1024;

    size_t klen = offsetof(kern_expression, u.load.desc) + sizeof(kern_varload_desc) * nload;
    std::vector<char> ebuf(klen + 16, 0);
    kern_expression *kexp = (kern_expression *)ebuf.data();
    kexp->len = klen;
    kexp->exptype = TypeOpCode__int4;
    kexp->opcode = FuncOpCode__LoadVars;
    kexp->nr_args = 0;
    kexp->u.load.depth = 0;
    kexp->u.load.nitems = nload;
    for (int i = 0; i < nload; i++)
    {
        kexp->u.load.desc[i].vl_resno = resnos[i];
        kexp->u.load.desc[i].vl_slot_id = i;
    }
    if (!ExecLoadVarsOuterHeap(kcxt, kexp, NULL, (const kern_data_store *)kds,
                               (const HeapTupleHeaderData *)htup))
        return -1;
    for (int i = 0; i < nload; i++)
    {
        xpu_datum_t *d = slots[i];
        nulls[i] = d->expr_ops == NULL;
        vals[i] = 0;
        offs[i] = 0;
        lens[i] = 0;
        if (nulls[i])
            continue;
        switch (types[i])
        {
            case TypeOpCode__bool: vals[i] = ((xpu_bool_t *)d)->value ? 1 : 0; break;
            case TypeOpCode__int1: vals[i] = ((xpu_int1_t *)d)->value; break;
            case TypeOpCode__int2: vals[i] = ((xpu_int2_t *)d)->value; break;
            case TypeOpCode__int4: vals[i] = ((xpu_int4_t *)d)->value; break;
            case TypeOpCode__int8: vals[i] = ((xpu_int8_t *)d)->value; break;
            case TypeOpCode__float4:
            {
                uint32_t b;
                memcpy(&b, &((xpu_float4_t *)d)->value, 4);
                vals[i] = b;
                break;
            }
            case TypeOpCode__float8: memcpy(&vals[i], &((xpu_float8_t *)d)->value, 8); break;
            case TypeOpCode__text:
            {
                xpu_text_t *t = (xpu_text_t *)d;
                offs[i] = (int64_t)((const char *)t->value - (const char *)htup);
                lens[i] = t->length;
                break;
            }
        }
    }
    return 0;
}

/*
 * The reference's device-function catalog as the reference's own opcode list
 * declares it (xpu_opcodes.h): device name -> opcode number, SQL name and
 * argument types (FUNC_OPCODE's FUNC_ARGS, e.g. "int4/int8").
 */
static std::string g_catalog;

extern "C" const char *
ref_func_catalog_json(void)
{
    g_catalog = "{";
    char buf[512];
#define FUNC_OPCODE(SQL,ARGS,FLAGS,DEV,COST,EXT)                               \
    snprintf(buf, sizeof(buf), "%s\"%s\": {\"opcode\": %d, \"sql\": \"%s\", \"args\": \"%s\"}", \
             g_catalog.size() > 1 ? ",\n " : "", #DEV, (int)FuncOpCode__##DEV, #SQL, #ARGS); \
    g_catalog += buf;
#define DEVONLY_FUNC_OPCODE(RET,DEV,ARGS,FLAGS,COST)                            \
    snprintf(buf, sizeof(buf), "%s\"%s\": {\"opcode\": %d, \"sql\": \"\", \"args\": \"%s\", \"ret\": \"%s\"}", \
             g_catalog.size() > 1 ? ",\n " : "", #DEV, (int)FuncOpCode__##DEV, #ARGS, #RET); \
    g_catalog += buf;
#include "xpu_opcodes.h"
#undef FUNC_OPCODE
#undef DEVONLY_FUNC_OPCODE
    g_catalog += "}\n";
    return g_catalog.c_str();
}

static const xpu_datum_operators *
ref_type_ops(int type_code)
{
    for (int i = 0; builtin_xpu_types_catalog[i].type_opcode != TypeOpCode__Invalid; i++)
        if ((int)builtin_xpu_types_catalog[i].type_opcode == type_code)
            return builtin_xpu_types_catalog[i].type_ops;
    return NULL;
}

static xpu_function_t
ref_func_dptr(int opcode)
{
    for (int i = 0; builtin_xpu_functions_catalog[i].func_opcode != FuncOpCode__Invalid; i++)
        if ((int)builtin_xpu_functions_catalog[i].func_opcode == opcode)
            return builtin_xpu_functions_catalog[i].func_dptr;
    return NULL;
}

/* the service's pointer resolution (gpu_service.c:3156-3255) on one tree */
static bool
ref_resolve(kern_expression *kexp)
{
    kexp->fn_dptr = ref_func_dptr(kexp->opcode);
    if (!kexp->fn_dptr)
        return false;
    kexp->expr_ops = ref_type_ops(kexp->exptype);
    if (kexp->opcode == FuncOpCode__CaseWhenExpr)
    {
        if (kexp->u.casewhen.case_comp &&
            !ref_resolve((kern_expression *)((char *)kexp + kexp->u.casewhen.case_comp)))
            return false;
        if (kexp->u.casewhen.case_else &&
            !ref_resolve((kern_expression *)((char *)kexp + kexp->u.casewhen.case_else)))
            return false;
    }
    if (kexp->opcode == FuncOpCode__HashedArrayOp)
        kexp->u.haop.elem_ops = ref_type_ops(kexp->u.haop.elem_type);
    kern_expression *karg = KEXP_FIRST_ARG(kexp);
    for (int i = 0; i < kexp->nr_args; i++, karg = KEXP_NEXT_ARG(karg))
        if (!ref_resolve(karg))
            return false;
    return true;
}

/*
 * Evaluates one kern_expression with the reference's own device functions
 * (EXEC_KERN_EXPRESSION, xpu_common.h:2694) over nslots variable slots.
 * Slot i has type code slot_types[i]; its value is the PostgreSQL datum
 * image at slot_data + slot_offs[i] (read with the type's
 * xpu_datum_heap_read), or SQL NULL if slot_offs[i] < 0.  On success
 * returns 0 with *res_type the result's type code and the result's datum
 * image (xpu_datum_write) in res_image, *res_len its length or -1 for NULL.
 * If the reference raises an error, returns 1 with errcode and message.
 * Returns -1 for a tree it cannot resolve.
 */
/* the session's xpu_encode_info for encname, resolved from the
 * reference's catalog as the service does (gpu_service.c:3284-3305);
 * encname NULL or empty leaves the session without one */
static bool
ref_session_encode(kern_session_info *session, size_t cap, const char *encname)
{
    if (!encname || !*encname)
        return true;
    size_t off = 2048;
    if (off + sizeof(xpu_encode_info) > cap)
        return false;
    xpu_encode_info *enc = (xpu_encode_info *)((char *)session + off);
    for (int i = 0; xpu_encode_catalog[i].enc_maxlen > 0; i++)
    {
        if (strcmp(xpu_encode_catalog[i].encname, encname) == 0)
        {
            *enc = xpu_encode_catalog[i];
            session->session_encode = off;
            return true;
        }
    }
    return false;
}

extern "C" int
ref_eval_expr(const void *kexp_blob, size_t kexp_len, int compare_nulls, const char *encname,
              int nslots, const int32_t *slot_types, const uint8_t *slot_data,
              const int64_t *slot_offs,
              int32_t *res_type, uint8_t *res_image, int32_t res_cap, int32_t *res_len,
              uint32_t *errcode, char *errmsg, int errlen)
{
    std::vector<uint64_t> ebuf((kexp_len + 7) / 8 + 1, 0);
    memcpy(ebuf.data(), kexp_blob, kexp_len);
    kern_expression *kexp = (kern_expression *)ebuf.data();
    if (!ref_resolve(kexp))
        return -1;
    std::vector<uint64_t> sbuf(4096 / 8, 0);      /* a session with no params */
    if (!ref_session_encode((kern_session_info *)sbuf.data(), 4096, encname))
        return -1;
    const size_t vl = 1 << 16;
    std::vector<uint64_t> kbuf((offsetof(kern_context, vlbuf) + vl) / 8 + 1, 0);
    kern_context *kcxt = (kern_context *)kbuf.data();
    kcxt->session = (kern_session_info *)sbuf.data();
    kcxt->kmode_compare_nulls = compare_nulls != 0;
    kcxt->vlpos = kcxt->vlbuf;
    kcxt->vlend = kcxt->vlbuf + vl;
    std::vector<std::vector<uint64_t>> store(nslots > 0 ? nslots : 1);
    std::vector<xpu_datum_t *> slots(nslots > 0 ? nslots : 1);
    std::vector<kern_varslot_desc> desc(nslots > 0 ? nslots : 1);
    for (int i = 0; i < nslots; i++)
    {
        const xpu_datum_operators *ops = ref_type_ops(slot_types[i]);
        if (!ops)
            return -1;
        store[i].assign(16, 0);
        slots[i] = (xpu_datum_t *)store[i].data();
        memset(&desc[i], 0, sizeof(kern_varslot_desc));
        desc[i].vs_type_code = (TypeOpCode)slot_types[i];
        desc[i].vs_typbyval = ops->xpu_type_byval;
        desc[i].vs_typalign = ops->xpu_type_align;
        desc[i].vs_typlen = ops->xpu_type_length;
        desc[i].vs_offset = -1;
        desc[i].vs_ops = ops;
        if (slot_offs[i] < 0)
            slots[i]->expr_ops = NULL;
        else if (!ops->xpu_datum_heap_read(kcxt, slot_data + slot_offs[i], slots[i]))
            return -1;
    }
    kcxt->kvars_slot = slots.data();
    kcxt->kvars_desc = desc.data();
    kcxt->kvars_nslots = nslots;
    kcxt->kvars_nrooms = nslots;
    std::vector<uint64_t> rbuf(16, 0);
    xpu_datum_t *r = (xpu_datum_t *)rbuf.data();
    *errcode = 0;
    if (!EXEC_KERN_EXPRESSION(kcxt, kexp, r))
    {
        *errcode = kcxt->errcode;
        snprintf(errmsg, errlen, "%s", kcxt->error_message ? kcxt->error_message : "");
        return 1;
    }
    if (XPU_DATUM_ISNULL(r))
    {
        *res_type = kexp->exptype;
        *res_len = -1;
        return 0;
    }
    const xpu_datum_operators *rops = r->expr_ops;
    *res_type = rops->xpu_type_code;
    if (rops->xpu_type_code == TypeOpCode__numeric)
    {
        /* NaN and the infinities: __decimal_to_varlena stores their header
         * and then overwrites it with SET_VARSIZE (xpu_numeric.h:289-300),
         * so the kind the reference computed is recorded in PostgreSQL's
         * own 6-byte encoding instead of the bytes it leaves behind */
        const xpu_numeric_t *nr = (const xpu_numeric_t *)r;
        if (nr->kind == XPU_NUMERIC_KIND__NAN || nr->kind == XPU_NUMERIC_KIND__POS_INF ||
            nr->kind == XPU_NUMERIC_KIND__NEG_INF)
        {
            uint32_t vl = 6 << 2;
            uint16_t h = nr->kind == XPU_NUMERIC_KIND__POS_INF ? 0xD000
                       : nr->kind == XPU_NUMERIC_KIND__NEG_INF ? 0xF000 : 0xC000;
            memcpy(res_image, &vl, 4);
            memcpy(res_image + 4, &h, 2);
            *res_len = 6;
            return 0;
        }
    }
    kern_colmeta cm;
    memset(&cm, 0, sizeof(cm));
    cm.attbyval = rops->xpu_type_byval;
    cm.attalign = rops->xpu_type_align;
    cm.attlen = rops->xpu_type_length;
    cm.atttypmod = -1;
    int n = rops->xpu_datum_write(kcxt, NULL, &cm, r);
    if (n < 0 || n > res_cap)
        return -1;
    rops->xpu_datum_write(kcxt, (char *)res_image, &cm, r);
    *res_len = n;
    return 0;
}

/* kcxt with nslots variable slots loaded from datum images (see
 * ref_eval_expr); false if a type is unknown */
struct RefSlots
{
    std::vector<uint64_t> kbuf, sbuf;
    std::vector<std::vector<uint64_t>> store;
    std::vector<xpu_datum_t *> slots;
    std::vector<kern_varslot_desc> desc;
    kern_context *kcxt = nullptr;

    bool load(int nslots, const int32_t *slot_types, const uint8_t *slot_data,
              const int64_t *slot_offs)
    {
        const size_t vl = 1 << 16;
        kbuf.assign((offsetof(kern_context, vlbuf) + vl) / 8 + 1, 0);
        sbuf.assign(4096 / 8, 0);
        kcxt = (kern_context *)kbuf.data();
        kcxt->session = (kern_session_info *)sbuf.data();
        kcxt->vlpos = kcxt->vlbuf;
        kcxt->vlend = kcxt->vlbuf + vl;
        int n = nslots > 0 ? nslots : 1;
        store.assign(n, std::vector<uint64_t>(16, 0));
        slots.assign(n, nullptr);
        desc.assign(n, kern_varslot_desc());
        for (int i = 0; i < nslots; i++)
        {
            const xpu_datum_operators *ops = ref_type_ops(slot_types[i]);
            if (!ops)
                return false;
            slots[i] = (xpu_datum_t *)store[i].data();
            memset(&desc[i], 0, sizeof(kern_varslot_desc));
            desc[i].vs_type_code = (TypeOpCode)slot_types[i];
            desc[i].vs_ops = ops;
            if (slot_offs[i] < 0)
                slots[i]->expr_ops = NULL;
            else if (!ops->xpu_datum_heap_read(kcxt, slot_data + slot_offs[i], slots[i]))
                return false;
        }
        kcxt->kvars_slot = slots.data();
        kcxt->kvars_desc = desc.data();
        kcxt->kvars_nslots = nslots;
        kcxt->kvars_nrooms = nslots;
        return true;
    }
};

/*
 * The reference's projection tuple former, kern_form_minimal_tuple
 * (xpu_common.cu:1096-1209), over nattrs values: attribute j is slot j, its
 * datum image at slot_data + slot_offs[j] (NULL if negative); kds_dst is the
 * destination's KDS_FORMAT_ROW header with its colmeta.  Writes the
 * kern_tupitem into titem (cap bytes) and returns its t_len, or -1.
 */
extern "C" int
ref_form_tuple(const void *kds_dst, int nattrs, const int32_t *slot_types,
               const uint8_t *slot_data, const int64_t *slot_offs, uint32_t rowid,
               void *titem, int cap)
{
    RefSlots rs;
    if (!rs.load(nattrs, slot_types, slot_data, slot_offs))
        return -1;
    size_t klen = offsetof(kern_expression, u.proj.slot_id) + 2 * nattrs + 8;
    std::vector<uint64_t> ebuf(klen / 8 + 2, 0);
    kern_expression *kexp = (kern_expression *)ebuf.data();
    kexp->len = klen;
    kexp->opcode = FuncOpCode__Projection;
    kexp->u.proj.nattrs = nattrs;
    for (int j = 0; j < nattrs; j++)
        kexp->u.proj.slot_id[j] = j;
    int sz = kern_form_minimal_tuple(rs.kcxt, kexp, (const kern_data_store *)kds_dst, NULL, rowid);
    if (sz < 0 || sz > cap)
        return -1;
    memset(titem, 0, cap);
    return kern_form_minimal_tuple(rs.kcxt, kexp, (const kern_data_store *)kds_dst,
                                   (kern_tupitem *)titem, rowid);
}

/*
 * Walks a KDS_FORMAT_HASH buffer the way the reference's probe does
 * (KDS_HASH_FIRST_ITEM / KDS_HASH_NEXT_ITEM, xpu_common.h:1424-1451): for
 * each hash value hashes[i], writes into rowids the row ids of the items on
 * its chain whose t.hash equals it, in chain order, at most cap in total;
 * counts[i] receives how many.  Returns the total, or -1 past cap.
 */
extern "C" int64_t
ref_hash_walk(const void *kds_hash, int nhashes, const uint32_t *hashes, uint32_t *counts,
              uint32_t *rowids, int64_t cap)
{
    const kern_data_store *kds = (const kern_data_store *)kds_hash;
    int64_t total = 0;
    for (int i = 0; i < nhashes; i++)
    {
        uint32_t c = 0;
        for (kern_hashitem *hitem = KDS_HASH_FIRST_ITEM(kds, hashes[i]);
             hitem != NULL;
             hitem = KDS_HASH_NEXT_ITEM(kds, hitem->next))
        {
            if (hitem->t.hash != hashes[i])
                continue;
            if (total >= cap)
                return -1;
            rowids[total++] = KERN_TUPITEM_GET_ROWID(&hitem->t);
            c++;
        }
        counts[i] = c;
    }
    return total;
}
