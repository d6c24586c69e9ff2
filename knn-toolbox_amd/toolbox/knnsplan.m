function plan = knnsplan(n, q, d, k)
% KNNSPLAN  plan a k-nearest-neighbour search of q queries against n corpus vectors of d
% dimensions.
%   plan = knnsplan(n, q, d, k)
%   plan.function: the entry point to call (@gpuknnHeap), plan.k, plan.streams.
%   Both entry points share the same fused GPU search and return identical results, and the
%   heap (threshold) selection is never slower on the sizes measured, so the plan is fixed
%   instead of being chosen by timing trial searches; n, q and d are accepted for
%   compatibility with the plans of the CUDA toolbox.
if k < 1 || k > 1024 || k ~= floor(k)
  error('knns:k', 'k must be an integer in [1, 1024]');
end
plan.function = @gpuknnHeap;
plan.k = k;
plan.streams = 1;
end
